// Batched flexible GMRES building blocks in complex128, one independent
// Krylov space per right-hand side, all right-hand sides advanced by the same
// launches ("block" Gram-Schmidt across RHS: one pass computes the projections
// of every RHS onto every basis vector).
//
// Orthogonalisation is classical Gram-Schmidt applied twice (CGS2), fused into
// three streaming passes over the basis:
//   pass 1  h1 = V^H w
//   pass 2  w' = w - V h1 (stored),  h2 = V^H w',  ||w'||^2
//   pass 3  v_{j+1} = (w' - V h2) / sqrt(||w'||^2 - ||h2||^2)
// Lazy mode (complex64 preconditioner): pass 1 also sums ||w||^2 and pass 2
// only forms w' and ||w'||^2 (no basis values held for a second product);
// the second projection h2 = V^H w' and its correction run only for RHS whose
// projection cancelled severely, ||w'|| < 1e-4 ||w|| (a Daniel-Gragg-Kaufman-
// Stewart test with a small threshold); otherwise w' is kept unnormalised and
// its scale 1/||w'|| is applied where the basis vector is read.
// The lazy mode can keep the basis in complex64 (TV = float): every product
// and sum is still formed in float64, only the stored basis values are
// rounded (about 6e-8 relative), which perturbs the Arnoldi relation and the
// orthogonality at that level while the outer iterate, the residual that
// each restart recomputes and the Hessenberg/Givens work stay complex128 —
// the compressed-basis GMRES of Aliaga, Anzt et al. (2022).  The stored
// complex64 w' is then the preconditioner's next input as it is.
// Partial sums are per (chunk, RHS) and reduced in a fixed order, so results
// are bit-reproducible run to run.  The small Hessenberg/Givens work runs on
// the device (one wavefront per RHS) so a whole iteration needs no host sync.
#include "hh_kernels.h"

namespace hh {

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// block-wide (256 threads = 4 waves) sum; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = sh[0] + sh[1] + sh[2] + sh[3];
  return r;
}

static __device__ __forceinline__ void chunk_range(int64_t N, int nchunk, int c, int64_t* a, int64_t* b) {
  const int64_t L = (N + nchunk - 1) / nchunk;
  *a = (int64_t)c * L;
  *b = min(N, *a + L);
}

// The lazy passes that read a stored w (k_dots, k_update_lazy) and those that
// form w = H z in place (k_dots_hz, k_update_lazy_hz) accumulate through these
// helpers, written as explicit fmas in a fixed order, so both forms give
// identical sums whatever the compiler contracts in either kernel.
// acc + conj(v) w
__device__ __forceinline__ cx<double> cmac_conj(cx<double> acc, cx<double> v, cx<double> w) {
  return mk<double>(fma(v.im, w.im, fma(v.re, w.re, acc.re)), fma(-v.im, w.re, fma(v.re, w.im, acc.im)));
}
// w - h v
__device__ __forceinline__ cx<double> csub_mul(cx<double> w, cx<double> h, cx<double> v) {
  return mk<double>(fma(h.im, v.im, fma(-h.re, v.re, w.re)), fma(-h.im, v.re, fma(-h.re, v.im, w.im)));
}
// nn + |w|^2
__device__ __forceinline__ double nrm_acc(double nn, cx<double> w) { return fma(w.im, w.im, fma(w.re, w.re, nn)); }

// pass 1: part[c][i][b] = sum_{p in chunk c} conj(V_i[p,b]) w[p,b], i <= j
// NORM: also partn[c][b] = sum ||w||^2 over the chunk.  only (may be null):
// skip RHS with only[b] == 0.
template <int NV, bool NORM, typename TV>
__global__ __launch_bounds__(256) void k_dots(KrylovWork kw, const cx<TV>* __restrict__ V,
                                              const cx<TV>* __restrict__ w, int j,
                                              const int* __restrict__ only) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (kw.done[b] || (only && !only[b])) return;
  __shared__ double sh[4];
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t stride = kw.N * kw.nrhs;
  const int64_t base = (int64_t)b * kw.N;
  cx<double> acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = mk<double>(0.0, 0.0);
  double nn = 0.0;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    const cx<double> wv = cvt<TV, double>(w[base + p]);
    if (NORM) nn = nrm_acc(nn, wv);
#pragma unroll
    for (int i = 0; i < NV; ++i)
      if (i <= j) acc[i] = cmac_conj(acc[i], cvt<TV, double>(V[i * stride + base + p]), wv);
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    if (i <= j) {  // (no early exit: keeps acc[] in registers)
      const double si = kw.lazy ? kw.vscale[i * kw.nrhs + b] : 1.0;
      const double re = block_sum(si * acc[i].re, sh);
      const double im = block_sum(si * acc[i].im, sh);
      if (threadIdx.x == 0) kw.part[((int64_t)c * (kw.kmax + 1) + i) * kw.nrhs + b] = mk<double>(re, im);
    }
  }
  if (NORM) {
    const double t = block_sum(nn, sh);
    if (threadIdx.x == 0) kw.partn[(int64_t)c * kw.nrhs + b] = t;
  }
}

// reduce part over chunks into h[i][b] (one wavefront per (i, b)); block
// i = j + 1 (launched in lazy mode) reduces partn into wn2[b]
__global__ __launch_bounds__(64) void k_reduce_dots(KrylovWork kw, cx<double>* __restrict__ h, int j) {
  const int i = blockIdx.x, b = blockIdx.y;
  if (kw.done[b] || i > j + 1) return;
  if (i == j + 1) {
    double nn = 0.0;
    for (int c = threadIdx.x; c < kw.nchunk; c += 64) nn += kw.partn[(int64_t)c * kw.nrhs + b];
    nn = wave_sum(nn);
    if (threadIdx.x == 0) kw.wn2[b] = nn;
    return;
  }
  double re = 0.0, im = 0.0;
  for (int c = threadIdx.x; c < kw.nchunk; c += 64) {
    const cx<double> v = kw.part[((int64_t)c * (kw.kmax + 1) + i) * kw.nrhs + b];
    re += v.re;
    im += v.im;
  }
  re = wave_sum(re);
  im = wave_sum(im);
  if (threadIdx.x == 0) {
    h[i * kw.nrhs + b] = mk<double>(re, im);
    if (kw.lazy) {
      const double si = kw.vscale[i * kw.nrhs + b];
      kw.h1s[i * kw.nrhs + b] = mk<double>(si * re, si * im);
    }
  }
}

// pass 2 (complex128 mode, normalised basis): w' = w - V h1 ; part = V^H w' ;
// partn = ||w'||^2.  h1 is read with uniform (scalar) loads.
template <int NV>
__global__ __launch_bounds__(256) void k_update_dots(KrylovWork kw, const cx<double>* __restrict__ V,
                                                     const cx<double>* w, cx<double>* wout, int j) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (kw.done[b]) return;
  __shared__ double sh[4];
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t stride = kw.N * kw.nrhs;
  const int64_t base = (int64_t)b * kw.N;
  const cx<double>* __restrict__ hc = kw.h1 + b;
  const int hstride = kw.nrhs;
  cx<double> acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = mk<double>(0.0, 0.0);
  double nn = 0.0;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    cx<double> vv[NV];
    cx<double> wv = w[base + p];
#pragma unroll
    for (int i = 0; i < NV; ++i)
      if (i <= j) {
        vv[i] = V[i * stride + base + p];
        wv = wv - hc[i * hstride] * vv[i];
      }
    wout[base + p] = wv;
    nn += cabs2(wv);
#pragma unroll
    for (int i = 0; i < NV; ++i)
      if (i <= j) acc[i] = acc[i] + cmulc(vv[i], wv);
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    if (i <= j) {
      const double re = block_sum(acc[i].re, sh);
      const double im = block_sum(acc[i].im, sh);
      if (threadIdx.x == 0) kw.part[((int64_t)c * (kw.kmax + 1) + i) * kw.nrhs + b] = mk<double>(re, im);
    }
  }
  const double t = block_sum(nn, sh);
  if (threadIdx.x == 0) kw.partn[(int64_t)c * kw.nrhs + b] = t;
}

// lazy pass 2: w' = w - V h1 (basis vectors carry their scale), partn = ||w'||^2
template <int NV>
__global__ __launch_bounds__(256) void k_update_lazy(KrylovWork kw, const cx<double>* __restrict__ V,
                                                     const cx<double>* __restrict__ w, cx<double>* __restrict__ wout,
                                                     int j) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (kw.done[b]) return;
  __shared__ double sh[4];
  __shared__ cx<double> hs[NV];
  if (threadIdx.x < NV) hs[threadIdx.x] = threadIdx.x <= j ? kw.h1s[threadIdx.x * kw.nrhs + b] : mk<double>(0.0, 0.0);
  __syncthreads();
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t stride = kw.N * kw.nrhs;
  const int64_t base = (int64_t)b * kw.N;
  double nn = 0.0;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    cx<double> wv = w[base + p];
#pragma unroll
    for (int i = 0; i < NV; ++i)
      if (i <= j) wv = csub_mul(wv, hs[i], V[i * stride + base + p]);
    wout[base + p] = wv;
    if (kw.stage) kw.stage[base + p] = mk<float>((float)wv.re, (float)wv.im);
    nn = nrm_acc(nn, wv);
  }
  const double t = block_sum(nn, sh);
  if (threadIdx.x == 0) kw.partn[(int64_t)c * kw.nrhs + b] = t;
}

// ---- lazy passes 1 and 2 with w = H z formed on the fly (complex64 z) ----
// w is never stored: each pass recomputes it from z (8 B per value) and the
// shared diagonal instead of reading a stored complex128 w (16 B) that a
// separate stencil launch wrote (16 B).  Blocks are ordered RHS-fastest so the
// blocks of one node chunk run together and share the diagonal in L2.

// (H z)_p on a 2-D or 3-D grid: 5/7-point operator with mirror (BC 2) or
// self (BC 1) ghosts and complex128 diagonal d, in float64 (apply5/apply7, as
// in the stencil kernels' APPLY mode, so the values equal a stored H z bit for
// bit).  zb: this RHS's block of z.  hz_load issues the loads, hz_eval forms
// the value, so a caller can put every load of an iteration in flight first.
template <int DIM>
struct HzOps {
  cx<float> c, x1l, x1r, x2l, x2r, x3l, x3r;
  cx<double> d;
};

template <int DIM>
__device__ __forceinline__ void hz_load(const Grid& g, const cx<float>* __restrict__ zb,
                                        const cx<double>* __restrict__ d, int ci, int cj, int ck, int64_t p,
                                        HzOps<DIM>& o) {
  const int n1 = g.n1, n2 = g.n2;
  const int64_t plane = (int64_t)n1 * n2, pl = DIM == 3 ? (int64_t)ck * plane : 0;
  const int64_t row = pl + (int64_t)cj * n1;
  o.c = zb[p];
  o.x1l = zb[row + lo_nb(ci, g.bc)];
  o.x1r = zb[row + hi_nb(ci, n1, g.bc)];
  o.x2l = zb[pl + (int64_t)lo_nb(cj, g.bc) * n1 + ci];
  o.x2r = zb[pl + (int64_t)hi_nb(cj, n2, g.bc) * n1 + ci];
  if (DIM == 3) {
    const int64_t col = (int64_t)cj * n1 + ci;
    o.x3l = zb[(int64_t)lo_nb(ck, g.bc) * plane + col];
    o.x3r = zb[(int64_t)hi_nb(ck, g.n3, g.bc) * plane + col];
  }
  o.d = d[p];
}

template <int DIM>
__device__ __forceinline__ cx<double> hz_eval(const Grid& g, const HzOps<DIM>& o) {
  const cx<double> xc = cvt<float, double>(o.c);
  const cx<double> x1l = cvt<float, double>(o.x1l), x1r = cvt<float, double>(o.x1r);
  const cx<double> x2l = cvt<float, double>(o.x2l), x2r = cvt<float, double>(o.x2r);
  if (DIM == 2) return apply5(xc, x1l, x1r, x2l, x2r, g.ih1, g.ih2, o.d);
  return apply7(xc, x1l, x1r, x2l, x2r, cvt<float, double>(o.x3l), cvt<float, double>(o.x3r), g.ih1, g.ih2, g.ih3,
                o.d);
}

// The same operands with the node and its in-plane neighbours from a sliding
// window of z in LDS (position q mod R, R a power of two): each of them lies
// in [p - n1, p + n1].  In 3-D the two neighbours a plane away are read from
// z itself (the window would have to span two planes).
template <int DIM>
__device__ __forceinline__ void hz_load_ring(const Grid& g, const cx<float>* zr, int rmask,
                                             const cx<float>* __restrict__ zb, const cx<double>* __restrict__ d,
                                             int ci, int cj, int ck, int64_t p, HzOps<DIM>& o) {
  const int n1 = g.n1, n2 = g.n2;
  const int64_t plane = (int64_t)n1 * n2, pl = DIM == 3 ? (int64_t)ck * plane : 0;
  const int64_t row = pl + (int64_t)cj * n1;
  o.c = zr[p & rmask];
  o.x1l = zr[(row + lo_nb(ci, g.bc)) & rmask];
  o.x1r = zr[(row + hi_nb(ci, n1, g.bc)) & rmask];
  o.x2l = zr[(pl + (int64_t)lo_nb(cj, g.bc) * n1 + ci) & rmask];
  o.x2r = zr[(pl + (int64_t)hi_nb(cj, n2, g.bc) * n1 + ci) & rmask];
  if (DIM == 3) {
    // (a thread whose nodes have ended walks past the last plane: clamped)
    const int kc = min(ck, g.n3 - 1);
    const int64_t col = (int64_t)cj * n1 + ci;
    o.x3l = zb[(int64_t)lo_nb(kc, g.bc) * plane + col];
    o.x3r = zb[(int64_t)hi_nb(kc, g.n3, g.bc) * plane + col];
  }
  o.d = d[p];
}

// The window: before the block's step over [pb, pb + S) (S = 256 U nodes) it
// holds z[q] for q in [pb - n1, pb + S + n1); during the step every thread
// loads its share of the next S values [pb + S + n1, pb + 2S + n1) from HBM
// and stores them after its own reads.  With R >= 2 n1 + 2 S a stored
// position and a read one differ by less than R, so they never share a
// slot: one barrier per step orders the stores before the next step's reads.
// Each z value crosses HBM once instead of up to three times (the row above
// has left the caches when the block comes back to it); in 3-D the values a
// plane away are still read from z (about one extra read per value).
template <int U>
__device__ __forceinline__ void zring_fill(cx<float>* zr, int rmask, const cx<float>* __restrict__ zb, int64_t N,
                                           int64_t lo, int64_t hi) {
  for (int64_t q = lo + threadIdx.x; q < hi; q += 256)
    if (q >= 0 && q < N) zr[q & rmask] = zb[q];
}

// node coordinates of p and their advance by `by` positions (a multiple of
// the block's 256 threads)
struct NodeWalk {
  int ci, cj, ck;
  __device__ __forceinline__ void init(const Grid& g, int64_t p) {
    ci = (int)(p % g.n1);
    const int64_t q = p / g.n1;
    cj = (int)(q % g.n2);
    ck = (int)(q / g.n2);
  }
  __device__ __forceinline__ void step(const Grid& g, int by) {
    ci += by;
    while (ci >= g.n1) {
      ci -= g.n1;
      if (++cj == g.n2) {
        cj = 0;
        ++ck;
      }
    }
  }
};

// Both passes walk their chunk with U nodes per thread per iteration (p and
// p + 256 for U = 2) and issue every load of an iteration — the stencil
// operands, the diagonal and the NV basis values — before using any, with
// no branch between them: basis vector i is read at min(i, j), the values
// past j (re-reads of v_j from cache) feed only accumulators that are never
// stored (pass 1) or are scaled by a zero coefficient (pass 2).  Written
// with a load-use chain per basis vector, the compiler waited on each load
// in turn and the passes were bound by load latency, not HBM bandwidth.  A
// node past the chunk end (U = 2) loads node p's values and contributes 0.
// U = 1 keeps the summation order of the stored-w kernels.  With the z
// window (RING) every thread runs every step of its block; a thread whose
// nodes have ended reads node p0's basis values and diagonal, window slots of
// no meaning (and in 3-D z a plane away at a clamped plane), and contributes
// nothing.  EARLY (all NV
// basis values held in registers at once) for NV <= 10; longer restarts
// read one basis value at a time within their register budget.

// pass 1: part = V^H (H z), partn = ||H z||^2
template <int NV, typename TV, int U, int DIM, bool RING>
__global__ __launch_bounds__(256) void k_dots_hz(KrylovWork kw, Grid g, const cx<TV>* __restrict__ V,
                                                 const cx<float>* __restrict__ Z, const cx<double>* __restrict__ d,
                                                 int j, int rmask) {
  constexpr bool EARLY = NV <= 10;
  constexpr int S = 256 * U;
  const int b = blockIdx.x, c = blockIdx.y;
  if (kw.done[b]) return;
  __shared__ double sh[4];
  extern __shared__ cx<float> zr[];
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t stride = kw.N * kw.nrhs;
  const int64_t base = (int64_t)b * kw.N;
  const cx<float>* __restrict__ zb = Z + base;
  cx<double> acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = mk<double>(0.0, 0.0);
  double nn = 0.0;
  const double zs = kw.zunscaled ? kw.vscale[j * kw.nrhs + b] : 1.0;
  NodeWalk nw[U];
#pragma unroll
  for (int u = 0; u < U; ++u) nw[u].init(g, p0 + threadIdx.x + 256 * u);
  if constexpr (RING) {
    zring_fill<U>(zr, rmask, zb, kw.N, p0 - g.n1, p0 + S + g.n1);
    __syncthreads();
  }
  // block-uniform steps (the window's barrier needs every thread in each step)
  for (int64_t pb = p0; pb < p1; pb += S) {
    const int64_t p = pb + threadIdx.x;
    if constexpr (!RING) {
      if (p >= p1) break;  // no barrier in this loop: a thread leaves when its nodes end
    }
    HzOps<DIM> ho[U];
    int64_t qq[U];
    bool ok[U];
    cx<TV> vv[EARLY ? U : 1][EARLY ? NV : 1];
    cx<float> zn[U];
    int64_t qn[U];
    if constexpr (RING) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        qn[u] = pb + S + g.n1 + threadIdx.x + 256 * u;
        zn[u] = qn[u] < kw.N ? zb[qn[u]] : mk<float>(0.f, 0.f);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ok[u] = p + 256 * u < p1;
      qq[u] = ok[u] ? p + 256 * u : (p < p1 ? p : p0);
      const int ci = ok[u] ? nw[u].ci : nw[0].ci, cj = ok[u] ? nw[u].cj : nw[0].cj;
      const int ck = ok[u] ? nw[u].ck : nw[0].ck;
      if constexpr (RING)
        hz_load_ring<DIM>(g, zr, rmask, zb, d, ci, cj, ck, qq[u], ho[u]);
      else
        hz_load<DIM>(g, zb, d, ci, cj, ck, qq[u], ho[u]);
    }
    if constexpr (EARLY) {
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int i = 0; i < NV; ++i) vv[u][i] = V[(int64_t)min(i, j) * stride + base + qq[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const cx<double> wv = ok[u] ? zs * hz_eval<DIM>(g, ho[u]) : mk<double>(0.0, 0.0);
      nn = nrm_acc(nn, wv);
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        if constexpr (EARLY)
          acc[i] = cmac_conj(acc[i], cvt<TV, double>(vv[u][i]), wv);
        else if (i <= j)  // (register budget: one basis value at a time)
          acc[i] = cmac_conj(acc[i], cvt<TV, double>(V[(int64_t)i * stride + base + qq[u]]), wv);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) nw[u].step(g, 256 * U);
    if constexpr (RING) {
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (qn[u] < kw.N) zr[qn[u] & rmask] = zn[u];
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    if (i <= j) {
      const double si = kw.vscale[i * kw.nrhs + b];
      const double re = block_sum(si * acc[i].re, sh);
      const double im = block_sum(si * acc[i].im, sh);
      if (threadIdx.x == 0) kw.part[((int64_t)c * (kw.kmax + 1) + i) * kw.nrhs + b] = mk<double>(re, im);
    }
  }
  const double t = block_sum(nn, sh);
  if (threadIdx.x == 0) kw.partn[(int64_t)c * kw.nrhs + b] = t;
}

// pass 2: w' = H z - V h1 (stored in the basis precision), partn = ||w'||^2
// of the unrounded w'
template <int NV, typename TV, int U, int DIM, bool RING>
__global__ __launch_bounds__(256) void k_update_lazy_hz(KrylovWork kw, Grid g, const cx<TV>* __restrict__ V,
                                                        const cx<float>* __restrict__ Z,
                                                        const cx<double>* __restrict__ d,
                                                        cx<TV>* __restrict__ wout, int j, int rmask) {
  constexpr bool EARLY = NV <= 10;
  constexpr int S = 256 * U;
  const int b = blockIdx.x, c = blockIdx.y;
  if (kw.done[b]) return;
  __shared__ double sh[4];
  __shared__ cx<double> hs[NV];
  extern __shared__ cx<float> zr[];
  // coefficients past j are 0: those terms subtract exact zeros
  if (threadIdx.x < NV) hs[threadIdx.x] = threadIdx.x <= j ? kw.h1s[threadIdx.x * kw.nrhs + b] : mk<double>(0.0, 0.0);
  __syncthreads();
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t stride = kw.N * kw.nrhs;
  const int64_t base = (int64_t)b * kw.N;
  const cx<float>* __restrict__ zb = Z + base;
  double nn = 0.0;
  const double zs = kw.zunscaled ? kw.vscale[j * kw.nrhs + b] : 1.0;
  NodeWalk nw[U];
#pragma unroll
  for (int u = 0; u < U; ++u) nw[u].init(g, p0 + threadIdx.x + 256 * u);
  if constexpr (RING) {
    zring_fill<U>(zr, rmask, zb, kw.N, p0 - g.n1, p0 + S + g.n1);
    __syncthreads();
  }
  // block-uniform steps (the window's barrier needs every thread in each step)
  for (int64_t pb = p0; pb < p1; pb += S) {
    const int64_t p = pb + threadIdx.x;
    if constexpr (!RING) {
      if (p >= p1) break;  // no barrier in this loop: a thread leaves when its nodes end
    }
    HzOps<DIM> ho[U];
    int64_t qq[U];
    bool ok[U];
    cx<TV> vv[EARLY ? U : 1][EARLY ? NV : 1];
    cx<float> zn[U];
    int64_t qn[U];
    if constexpr (RING) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        qn[u] = pb + S + g.n1 + threadIdx.x + 256 * u;
        zn[u] = qn[u] < kw.N ? zb[qn[u]] : mk<float>(0.f, 0.f);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ok[u] = p + 256 * u < p1;
      qq[u] = ok[u] ? p + 256 * u : (p < p1 ? p : p0);
      const int ci = ok[u] ? nw[u].ci : nw[0].ci, cj = ok[u] ? nw[u].cj : nw[0].cj;
      const int ck = ok[u] ? nw[u].ck : nw[0].ck;
      if constexpr (RING)
        hz_load_ring<DIM>(g, zr, rmask, zb, d, ci, cj, ck, qq[u], ho[u]);
      else
        hz_load<DIM>(g, zb, d, ci, cj, ck, qq[u], ho[u]);
    }
    if constexpr (EARLY) {
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int i = 0; i < NV; ++i) vv[u][i] = V[(int64_t)min(i, j) * stride + base + qq[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cx<double> wv = zs * hz_eval<DIM>(g, ho[u]);
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        if constexpr (EARLY)
          wv = csub_mul(wv, hs[i], cvt<TV, double>(vv[u][i]));
        else if (i <= j)
          wv = csub_mul(wv, hs[i], cvt<TV, double>(V[(int64_t)i * stride + base + qq[u]]));
      }
      if (ok[u]) {
        wout[base + qq[u]] = cvt<double, TV>(wv);
        if (kw.stage) kw.stage[base + qq[u]] = mk<float>((float)wv.re, (float)wv.im);
        nn = nrm_acc(nn, wv);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) nw[u].step(g, 256 * U);
    if constexpr (RING) {
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (qn[u] < kw.N) zr[qn[u] & rmask] = zn[u];
      __syncthreads();
    }
  }
  const double t = block_sum(nn, sh);
  if (threadIdx.x == 0) kw.partn[(int64_t)c * kw.nrhs + b] = t;
}

// ---- 3-D: the same two passes marching z over in-plane columns ----
// Chunk c of a RHS is the in-plane segment c % nseg3 (C3_LQ consecutive
// nodes of a plane, two per thread) over the run c / nseg3 of the planes.
// LDS holds z on the windows [q0 - n1, q0 + C3_LQ + n1) of four planes:
// planes k - 1, k, k + 1 give H z on plane k (the in-plane neighbours and the
// values a plane away), plane k + 2 is loaded during step k and stored after
// it, one barrier per plane.  In the linear walk the values a plane away
// were fetched again from HBM (the plane had left L2 by the time a block
// came back to it): 1.2-1.45x the algorithmic bytes; here each z value is
// loaded once per block and the window's halo rows are shared in L2 with
// the neighbouring segments' blocks, which run beside it.
constexpr int C3_LQ = 512;

struct Col3 {
  int q0, q1, w0, w1, k0, k1;  // segment [q0, q1), window [w0, w1), planes [k0, k1)
};
__device__ __forceinline__ Col3 col3_of(const Grid& g, int c, int nseg) {
  Col3 r;
  const int plane = g.n1 * g.n2;
  const int sg = c % nseg, zr = c / nseg, nzr = max(1, (int)(gridDim.y / nseg));
  r.q0 = sg * C3_LQ;
  r.q1 = min(r.q0 + C3_LQ, plane);
  r.w0 = max(r.q0 - g.n1, 0);
  r.w1 = min(r.q1 + g.n1, plane);
  r.k0 = (int)((int64_t)g.n3 * zr / nzr);
  r.k1 = (int)((int64_t)g.n3 * (zr + 1) / nzr);
  return r;
}

// window of plane k (in [0, n3)) into its slot k mod 4 (slots of WS values)
template <int M>
__device__ __forceinline__ void col3_fill(cx<float>* zw, const cx<float>* __restrict__ zb, const Col3& cb,
                                          int64_t plane, int k) {
  constexpr int WS = 256 * M;
  cx<float>* dst = zw + (k & 3) * WS;
  const cx<float>* src = zb + (int64_t)k * plane + cb.w0;
  for (int e = threadIdx.x; e < cb.w1 - cb.w0; e += 256) dst[e] = src[e];
}

// H z operands of in-plane node q = cj n1 + ci on plane k from the windows
template <int M>
__device__ __forceinline__ void col3_ops(const Grid& g, const cx<float>* zw, const Col3& cb,
                                         const cx<double>* __restrict__ d, int k, int q, int ci, int cj,
                                         HzOps<3>& o) {
  constexpr int WS = 256 * M;
  const int n1 = g.n1;
  const cx<float>* wk = zw + (k & 3) * WS;
  const int row = cj * n1;
  o.c = wk[q - cb.w0];
  o.x1l = wk[row + lo_nb(ci, g.bc) - cb.w0];
  o.x1r = wk[row + hi_nb(ci, n1, g.bc) - cb.w0];
  o.x2l = wk[lo_nb(cj, g.bc) * n1 + ci - cb.w0];
  o.x2r = wk[hi_nb(cj, g.n2, g.bc) * n1 + ci - cb.w0];
  o.x3l = zw[(lo_nb(k, g.bc) & 3) * WS + q - cb.w0];
  o.x3r = zw[(hi_nb(k, g.n3, g.bc) & 3) * WS + q - cb.w0];
  o.d = d[(int64_t)k * n1 * g.n2 + q];
}

// pass 1 (3-D columns): part = V^H (H z), partn = ||H z||^2
template <int NV, typename TV, int M>
__global__ __launch_bounds__(256) void k_dots_hz3c(KrylovWork kw, Grid g, const cx<TV>* __restrict__ V,
                                                   const cx<float>* __restrict__ Z, const cx<double>* __restrict__ d,
                                                   int j) {
  constexpr bool EARLY = NV <= 10;
  constexpr int PT = C3_LQ / 256;  // in-plane nodes per thread (two, taken in turn)
  const int b = blockIdx.x, c = blockIdx.y;
  if (kw.done[b]) return;
  __shared__ double sh[4];
  extern __shared__ cx<float> zw[];
  const Col3 cb = col3_of(g, c, kw.nseg3);
  const int64_t plane = (int64_t)g.n1 * g.n2;
  const int64_t stride = kw.N * kw.nrhs;
  const int64_t base = (int64_t)b * kw.N;
  const cx<float>* __restrict__ zb = Z + base;
  cx<double> acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = mk<double>(0.0, 0.0);
  double nn = 0.0;
  const double zs = kw.zunscaled ? kw.vscale[j * kw.nrhs + b] : 1.0;
  int qv[PT], ci[PT], cj[PT];
  bool ok[PT];
#pragma unroll
  for (int t = 0; t < PT; ++t) {
    const int q = cb.q0 + threadIdx.x + 256 * t;
    ok[t] = q < cb.q1;
    qv[t] = ok[t] ? q : cb.q0;  // a node past the segment reads node q0 and contributes 0
    ci[t] = qv[t] % g.n1;
    cj[t] = qv[t] / g.n1;
  }
  for (int k = max(cb.k0 - 1, 0); k <= min(cb.k0 + 1, g.n3 - 1); ++k) col3_fill<M>(zw, zb, cb, plane, k);
  __syncthreads();
  for (int k = cb.k0; k < cb.k1; ++k) {
    // the window of plane k + 2, stored after this step's reads
    const int kp = k + 2;
    const bool pf = kp < g.n3 && kp <= cb.k1;
    cx<float> zn[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int e = threadIdx.x + 256 * m;
      zn[m] = pf && e < cb.w1 - cb.w0 ? zb[(int64_t)kp * plane + cb.w0 + e] : mk<float>(0.f, 0.f);
    }
    // the thread's nodes one at a time (each with all its loads issued first)
#pragma unroll 1
    for (int t = 0; t < PT; ++t) {
      const int q = t == 0 ? qv[0] : qv[1], qi = t == 0 ? ci[0] : ci[1], qj = t == 0 ? cj[0] : cj[1];
      const bool qok = t == 0 ? ok[0] : ok[1];
      const int64_t p = k * plane + q;
      HzOps<3> ho;
      cx<TV> vv[EARLY ? NV : 1];
      col3_ops<M>(g, zw, cb, d, k, q, qi, qj, ho);
      if constexpr (EARLY) {
#pragma unroll
        for (int i = 0; i < NV; ++i) vv[i] = V[(int64_t)min(i, j) * stride + base + p];
      }
      const cx<double> wv = qok ? zs * hz_eval<3>(g, ho) : mk<double>(0.0, 0.0);
      nn = nrm_acc(nn, wv);
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        if constexpr (EARLY)
          acc[i] = cmac_conj(acc[i], cvt<TV, double>(vv[i]), wv);
        else if (i <= j)
          acc[i] = cmac_conj(acc[i], cvt<TV, double>(V[(int64_t)i * stride + base + p]), wv);
      }
    }
    if (pf) {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int e = threadIdx.x + 256 * m;
        if (e < cb.w1 - cb.w0) zw[(kp & 3) * (256 * M) + e] = zn[m];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    if (i <= j) {
      const double si = kw.vscale[i * kw.nrhs + b];
      const double re = block_sum(si * acc[i].re, sh);
      const double im = block_sum(si * acc[i].im, sh);
      if (threadIdx.x == 0) kw.part[((int64_t)c * (kw.kmax + 1) + i) * kw.nrhs + b] = mk<double>(re, im);
    }
  }
  const double t = block_sum(nn, sh);
  if (threadIdx.x == 0) kw.partn[(int64_t)c * kw.nrhs + b] = t;
}

// pass 2 (3-D columns): w' = H z - V h1 (stored in the basis precision),
// partn = ||w'||^2 of the unrounded w'
template <int NV, typename TV, int M>
__global__ __launch_bounds__(256) void k_update_lazy_hz3c(KrylovWork kw, Grid g, const cx<TV>* __restrict__ V,
                                                          const cx<float>* __restrict__ Z,
                                                          const cx<double>* __restrict__ d,
                                                          cx<TV>* __restrict__ wout, int j) {
  constexpr bool EARLY = NV <= 10;
  constexpr int PT = C3_LQ / 256;
  const int b = blockIdx.x, c = blockIdx.y;
  if (kw.done[b]) return;
  __shared__ double sh[4];
  __shared__ cx<double> hs[NV];
  extern __shared__ cx<float> zw[];
  if (threadIdx.x < NV) hs[threadIdx.x] = threadIdx.x <= j ? kw.h1s[threadIdx.x * kw.nrhs + b] : mk<double>(0.0, 0.0);
  const Col3 cb = col3_of(g, c, kw.nseg3);
  const int64_t plane = (int64_t)g.n1 * g.n2;
  const int64_t stride = kw.N * kw.nrhs;
  const int64_t base = (int64_t)b * kw.N;
  const cx<float>* __restrict__ zb = Z + base;
  double nn = 0.0;
  const double zs = kw.zunscaled ? kw.vscale[j * kw.nrhs + b] : 1.0;
  int qv[PT], ci[PT], cj[PT];
  bool ok[PT];
#pragma unroll
  for (int t = 0; t < PT; ++t) {
    const int q = cb.q0 + threadIdx.x + 256 * t;
    ok[t] = q < cb.q1;
    qv[t] = ok[t] ? q : cb.q0;
    ci[t] = qv[t] % g.n1;
    cj[t] = qv[t] / g.n1;
  }
  for (int k = max(cb.k0 - 1, 0); k <= min(cb.k0 + 1, g.n3 - 1); ++k) col3_fill<M>(zw, zb, cb, plane, k);
  __syncthreads();
  for (int k = cb.k0; k < cb.k1; ++k) {
    const int kp = k + 2;
    const bool pf = kp < g.n3 && kp <= cb.k1;
    cx<float> zn[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int e = threadIdx.x + 256 * m;
      zn[m] = pf && e < cb.w1 - cb.w0 ? zb[(int64_t)kp * plane + cb.w0 + e] : mk<float>(0.f, 0.f);
    }
#pragma unroll 1
    for (int t = 0; t < PT; ++t) {
      const int q = t == 0 ? qv[0] : qv[1], qi = t == 0 ? ci[0] : ci[1], qj = t == 0 ? cj[0] : cj[1];
      const bool qok = t == 0 ? ok[0] : ok[1];
      const int64_t p = k * plane + q;
      HzOps<3> ho;
      cx<TV> vv[EARLY ? NV : 1];
      col3_ops<M>(g, zw, cb, d, k, q, qi, qj, ho);
      if constexpr (EARLY) {
#pragma unroll
        for (int i = 0; i < NV; ++i) vv[i] = V[(int64_t)min(i, j) * stride + base + p];
      }
      cx<double> wv = zs * hz_eval<3>(g, ho);
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        if constexpr (EARLY)
          wv = csub_mul(wv, hs[i], cvt<TV, double>(vv[i]));
        else if (i <= j)
          wv = csub_mul(wv, hs[i], cvt<TV, double>(V[(int64_t)i * stride + base + p]));
      }
      if (qok) {
        wout[base + p] = cvt<double, TV>(wv);
        if (kw.stage) kw.stage[base + p] = mk<float>((float)wv.re, (float)wv.im);
        nn = nrm_acc(nn, wv);
      }
    }
    if (pf) {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int e = threadIdx.x + 256 * m;
        if (e < cb.w1 - cb.w0) zw[(kp & 3) * (256 * M) + e] = zn[m];
      }
    }
    __syncthreads();
  }
  const double t = block_sum(nn, sh);
  if (threadIdx.x == 0) kw.partn[(int64_t)c * kw.nrhs + b] = t;
}

bool col3_ok(const Grid& g) { return g.dim == 3 && g.n1 <= 768 && (int64_t)g.n1 * g.n2 < (1LL << 30) && g.n3 >= 2; }
int col3_segments(const Grid& g) { return (int)(((int64_t)g.n1 * g.n2 + C3_LQ - 1) / C3_LQ); }

// lazy mode: fix[b] = 1 where the projection cancelled severely,
// ||w'||^2 < 1e-8 ||w||^2 (one wavefront per RHS)
__global__ __launch_bounds__(64) void k_fix_flags(KrylovWork kw) {
  const int b = blockIdx.x;
  if (kw.done[b]) return;
  double nn = 0.0;
  for (int c = threadIdx.x; c < kw.nchunk; c += 64) nn += kw.partn[(int64_t)c * kw.nrhs + b];
  nn = wave_sum(nn);
  if (threadIdx.x == 0) kw.fix[b] = !(nn >= 1e-8 * kw.wn2[b]);
}

// Reduce pass-2 sums, form column j of H, apply/update Givens rotations and
// the residual estimate; one wavefront per RHS.
__global__ __launch_bounds__(64) void k_hess_givens(KrylovWork kw, int j, double tol) {
  const int b = blockIdx.x;
  if (kw.done[b]) return;
  const int lane = threadIdx.x;
  __shared__ cx<double> col[72];
  // h2 and ||w'||^2 (lazy mode: h2 exists only where fix[b] was raised)
  const bool have_h2 = !kw.lazy || kw.fix[b];
  double s2 = 0.0;
  for (int i = 0; i <= j; ++i) {
    if (!have_h2) {
      if (lane == 0) col[i] = kw.h1[i * kw.nrhs + b];
      continue;
    }
    double re = 0.0, im = 0.0;
    for (int c = lane; c < kw.nchunk; c += 64) {
      const cx<double> v = kw.part[((int64_t)c * (kw.kmax + 1) + i) * kw.nrhs + b];
      re += v.re;
      im += v.im;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) {
      const cx<double> h2 = mk<double>(re, im);
      kw.h2[i * kw.nrhs + b] = h2;
      col[i] = kw.h1[i * kw.nrhs + b];
      s2 += cabs2(h2);
    }
  }
  double nn = 0.0;
  for (int c = lane; c < kw.nchunk; c += 64) nn += kw.partn[(int64_t)c * kw.nrhs + b];
  nn = wave_sum(nn);
  if (lane != 0) return;
  // the Hessenberg column describes exactly the vector that is kept
  const bool use_h2 = have_h2;
  if (use_h2)
    for (int i = 0; i <= j; ++i) col[i] = col[i] + kw.h2[i * kw.nrhs + b];
  const double hn2 = use_h2 ? nn - s2 : nn;
  const double hn = hn2 > 0.0 ? sqrt(hn2) : 0.0;
  kw.nrm[b] = hn > 0.0 ? 1.0 / hn : 0.0;
  if (kw.lazy) {
    kw.fix[b] = use_h2;
    kw.vscale[(j + 1) * kw.nrhs + b] = use_h2 ? 1.0 : kw.nrm[b];
  }
  col[j + 1] = mk<double>(hn, 0.0);
  const int K = kw.kmax;
  cx<double>* Hb = kw.H + (int64_t)b * (K + 1) * K;
  cx<double>* csb = kw.cs + (int64_t)b * K;
  cx<double>* snb = kw.sn + (int64_t)b * K;
  cx<double>* gb = kw.gv + (int64_t)b * (K + 1);
  for (int i = 0; i < j; ++i) {
    const cx<double> a = col[i], c2 = col[i + 1];
    const cx<double> t = csb[i] * a + snb[i] * c2;
    col[i + 1] = cmulc(snb[i], a) * mk<double>(-1.0, 0.0) + cmulc(csb[i], c2);
    col[i] = t;
  }
  const cx<double> a = col[j], bb = col[j + 1];
  const double den = sqrt(cabs2(a) + cabs2(bb));
  cx<double> cj, sj;
  if (den > 0.0) {
    cj = mk<double>(a.re / den, -a.im / den);
    sj = mk<double>(bb.re / den, -bb.im / den);
  } else {
    cj = mk<double>(1.0, 0.0);
    sj = mk<double>(0.0, 0.0);
  }
  csb[j] = cj;
  snb[j] = sj;
  col[j] = mk<double>(den, 0.0);
  for (int i = 0; i <= j; ++i) Hb[i + (int64_t)j * (K + 1)] = col[i];
  const cx<double> g = gb[j];
  gb[j + 1] = cmulc(sj, g) * mk<double>(-1.0, 0.0);
  gb[j] = cj * g;
  const double est = sqrt(cabs2(gb[j + 1])) / kw.bnorm[b];
  if (den > 0.0) {
    kw.kact[b] = j + 1;
    kw.est[b] = est;
  }
  if (!(den > 0.0) || est <= tol || !(est == est) || j + 1 >= K) kw.done[b] = 1;
}

// pass 3: v_{j+1} = (w' - V h2) * nrm
template <int NV, typename TV>
__global__ __launch_bounds__(256) void k_final(KrylovWork kw, const cx<TV>* __restrict__ V,
                                               const cx<TV>* w, cx<TV>* vout,
                                               int j, const int* __restrict__ skip_before) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (skip_before[b] || (kw.lazy && !kw.fix[b])) return;
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t stride = kw.N * kw.nrhs;
  const int64_t base = (int64_t)b * kw.N;
  cx<double> hc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double si = (kw.lazy && i <= j) ? kw.vscale[i * kw.nrhs + b] : 1.0;
    hc[i] = i <= j ? si * kw.h2[i * kw.nrhs + b] : mk<double>(0.0, 0.0);
  }
  const double s = kw.nrm[b];
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    cx<double> wv = cvt<TV, double>(w[base + p]);
#pragma unroll
    for (int i = 0; i < NV; ++i)
      if (i <= j) wv = wv - hc[i] * cvt<TV, double>(V[i * stride + base + p]);
    vout[base + p] = cvt<double, TV>(s * wv);
  }
}

// ||v_b||^2 partials and reduction
__global__ __launch_bounds__(256) void k_norm_part(KrylovWork kw, const cx<double>* __restrict__ v,
                                                   const int* __restrict__ skip) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (skip && skip[b]) return;
  __shared__ double sh[4];
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t base = (int64_t)b * kw.N;
  double nn = 0.0;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) nn += cabs2(v[base + p]);
  const double t = block_sum(nn, sh);
  if (threadIdx.x == 0) kw.partn[(int64_t)c * kw.nrhs + b] = t;
}

__global__ __launch_bounds__(64) void k_norm_reduce(KrylovWork kw, double* __restrict__ out,
                                                    const int* __restrict__ skip) {
  const int b = blockIdx.x;
  if (skip && skip[b]) return;
  double nn = 0.0;
  for (int c = threadIdx.x; c < kw.nchunk; c += 64) nn += kw.partn[(int64_t)c * kw.nrhs + b];
  nn = wave_sum(nn);
  if (threadIdx.x == 0) out[b] = nn;
}

// cycle start: beta = ||r||, g = [beta, 0...], flags
__global__ __launch_bounds__(64) void k_start_flags(KrylovWork kw, const double* __restrict__ n2, double tol,
                                                    const int* __restrict__ conv) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= kw.nrhs) return;
  kw.kact[b] = 0;
  const double bn = kw.bnorm[b];
  const double beta = conv[b] ? 0.0 : sqrt(n2[b]);
  int d = conv[b] || !(bn > 0.0) || !(beta == beta) || beta <= tol * bn;
  kw.done[b] = d;
  kw.nrm[b] = (!d && beta > 0.0) ? 1.0 / beta : 0.0;
  kw.est[b] = bn > 0.0 ? beta / bn : 0.0;
  cx<double>* gb = kw.gv + (int64_t)b * (kw.kmax + 1);
  gb[0] = mk<double>(beta, 0.0);
  if (kw.lazy) kw.vscale[b] = 1.0;
}

template <typename TV>
__global__ __launch_bounds__(256) void k_scale(KrylovWork kw, const cx<double>* __restrict__ r,
                                               cx<TV>* __restrict__ v) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (kw.done[b]) return;
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t base = (int64_t)b * kw.N;
  const double s = kw.nrm[b];
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) v[base + p] = cvt<double, TV>(s * r[base + p]);
}

// y = R^-1 g (back substitution), one thread per RHS
__global__ __launch_bounds__(64) void k_backsolve(KrylovWork kw) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= kw.nrhs) return;
  const int k = kw.kact[b];
  const int K = kw.kmax;
  const cx<double>* Hb = kw.H + (int64_t)b * (K + 1) * K;
  const cx<double>* gb = kw.gv + (int64_t)b * (K + 1);
  cx<double>* yb = kw.y + (int64_t)b * K;
  for (int i = k - 1; i >= 0; --i) {
    cx<double> s = gb[i];
    for (int l = i + 1; l < k; ++l) s = s - Hb[i + (int64_t)l * (K + 1)] * yb[l];
    yb[i] = s * cinv(Hb[i + (int64_t)i * (K + 1)]);
  }
  // x += sum y_i z_i with z_i = vscale_i Z_i
  if (kw.zunscaled)
    for (int i = 0; i < k; ++i) yb[i] = kw.vscale[i * kw.nrhs + b] * yb[i];
}

template <typename TZ>
__global__ __launch_bounds__(256) void k_xupdate(KrylovWork kw, const cx<TZ>* __restrict__ Z,
                                                 cx<double>* __restrict__ x) {
  const int c = blockIdx.x, b = blockIdx.y;
  const int k = kw.kact[b];
  if (k == 0) return;
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t stride = kw.N * kw.nrhs;
  const int64_t base = (int64_t)b * kw.N;
  const cx<double>* yb = kw.y + (int64_t)b * kw.kmax;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    cx<double> acc = x[base + p];
    for (int i = 0; i < k; ++i) acc = acc + yb[i] * cvt<TZ, double>(Z[i * stride + base + p]);
    x[base + p] = acc;
  }
}

__global__ __launch_bounds__(64) void k_div_bnorm(KrylovWork kw, const double* __restrict__ n2,
                                                  double* __restrict__ res) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= kw.nrhs) return;
  const double bn = kw.bnorm[b];
  res[b] = bn > 0.0 ? sqrt(n2[b]) / bn : 0.0;
}

template <typename TV>
__global__ __launch_bounds__(256) void k_cast_scaled(const cx<TV>* __restrict__ v, const double* __restrict__ sc,
                                                     cx<float>* __restrict__ o, int64_t N, int nrhs) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= N * nrhs) return;
  const double s = sc[k / N];
  const cx<double> a = cvt<TV, double>(v[k]);
  o[k] = mk<float>((float)(s * a.re), (float)(s * a.im));
}

// stage = (complex64) vout for RHS whose basis vector pass 3 rewrote
__global__ __launch_bounds__(256) void k_restage(KrylovWork kw, const cx<double>* __restrict__ vout,
                                                 const int* __restrict__ skip_before) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (skip_before[b] || !kw.fix[b]) return;
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t base = (int64_t)b * kw.N;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    const cx<double> v = vout[base + p];
    kw.stage[base + p] = mk<float>((float)v.re, (float)v.im);
  }
}

// wv *= sc[b] per RHS
__global__ __launch_bounds__(256) void k_scale_rhs(KrylovWork kw, cx<double>* __restrict__ wv,
                                                   const double* __restrict__ sc) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (kw.done[b]) return;
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t base = (int64_t)b * kw.N;
  const double s = sc[b];
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) wv[base + p] = s * wv[base + p];
}

// ------------------------------------------------------------------ host side
static inline dim3 cgrid(const KrylovWork& w) { return dim3(w.nchunk, w.nrhs); }

template <typename TV>
void kr_cast_scaled(const KrylovWork& w, const cx<TV>* Vj, const double* scale, cx<float>* b0, hipStream_t s) {
  const int64_t n = w.N * (int64_t)w.nrhs;
  hipLaunchKernelGGL((k_cast_scaled<TV>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Vj, scale, b0, w.N,
                     w.nrhs);
}
template void kr_cast_scaled<double>(const KrylovWork&, const cx<double>*, const double*, cx<float>*, hipStream_t);
template void kr_cast_scaled<float>(const KrylovWork&, const cx<float>*, const double*, cx<float>*, hipStream_t);

void kr_scale_rhs(const KrylovWork& w, cx<double>* wv, const double* scale, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_rhs, cgrid(w), dim3(256), 0, s, w, wv, scale);
}

void kr_norm2(const KrylovWork& w, const cx<double>* v, double* out, const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(k_norm_part, cgrid(w), dim3(256), 0, s, w, v, skip);
  hipLaunchKernelGGL(k_norm_reduce, dim3(w.nrhs), dim3(64), 0, s, w, out, skip);
}

// r: residual block, conv: per-RHS converged flags (device).  Uses w.nrm as scratch.
template <typename TV>
void kr_start_impl(const KrylovWork& w, const cx<double>* r, cx<TV>* V0, double tol, const int* conv,
                   double* n2tmp, hipStream_t s) {
  kr_norm2(w, r, n2tmp, conv, s);
  hipLaunchKernelGGL(k_start_flags, dim3((w.nrhs + 63) / 64), dim3(64), 0, s, w, n2tmp, tol, conv);
  hipLaunchKernelGGL((k_scale<TV>), cgrid(w), dim3(256), 0, s, w, r, V0);
}
template void kr_start_impl<double>(const KrylovWork&, const cx<double>*, cx<double>*, double, const int*, double*,
                                    hipStream_t);
template void kr_start_impl<float>(const KrylovWork&, const cx<double>*, cx<float>*, double, const int*, double*,
                                   hipStream_t);

// kt (optional) times pass 1 / pass 2; algorithmic bytes per active RHS:
// 16 N (j + 2) for pass 1 (reads v_0..v_j and w), 16 N (j + 3) for pass 2
// (also writes w')
//
// hz (optional, lazy mode): w = H z is formed inside passes 1 and 2 from the
// complex64 z; bytes per active RHS sv N (j + 1) + 8 N and sv N (j + 2) + 8 N
// (sv = 16 or 8 B per stored basis value), plus 16 N for the shared
// diagonal.  A staged input adds 8 N to pass 2.  A complex64 basis (TV =
// float) needs hz.
template <int NV, typename TV>
static void arnoldi_nv(const KrylovWork& w, const cx<TV>* V, cx<double>* wv, int j, double tol,
                       int* skip_copy, KTimer* kt, int active, const HzArgs* hz, hipStream_t s) {
  const int64_t stride = w.N * w.nrhs;
  const bool t1 = kt && kt->match(KT_DOTS, 0), t2 = kt && kt->match(KT_UPDATE, 0);
  cx<TV>* vnext = (cx<TV>*)(V + (j + 1) * stride);
  const double Nf = (double)w.N, sv = (double)sizeof(cx<TV>);
  if (t1) kt->begin(s);
  // nodes per thread per iteration: 2 for a complex64 basis up to NV = 10
  constexpr int U = std::is_same<TV, float>::value && NV <= 10 ? 2 : 1;
  // z (2-D; in 3-D the in-plane operands) through the LDS window when
  // 2 n1 + 2 S nodes (S = 256 U) fit in a power-of-two ring of at most 8192
  // complex64 values (64 KiB)
  int rnodes = 0;
  if (hz) {
    const int64_t need = 2 * (int64_t)hz->g.n1 + 2 * 256 * U;
    int64_t r = 1;
    while (r < need) r <<= 1;
    if (r <= 8192) rnodes = (int)r;
  }
  const size_t rbytes = sizeof(cx<float>) * (size_t)rnodes;
  // 3-D: z march over in-plane segments, four plane windows of 1024 or 2048 values
  const bool c3 = hz && hz->g.dim == 3 && w.nseg3 > 0;
  const int m3 = c3 && C3_LQ + 2 * hz->g.n1 > 1024 ? 8 : 4;
  const size_t c3bytes = sizeof(cx<float>) * 4 * 256 * (size_t)m3;
  if (c3) {
    if (m3 == 4)
      hipLaunchKernelGGL((k_dots_hz3c<NV, TV, 4>), dim3(w.nrhs, w.nchunk), dim3(256), c3bytes, s, w, hz->g, V,
                         hz->z, hz->d, j);
    else
      hipLaunchKernelGGL((k_dots_hz3c<NV, TV, 8>), dim3(w.nrhs, w.nchunk), dim3(256), c3bytes, s, w, hz->g, V,
                         hz->z, hz->d, j);
  } else if (hz) {
    if (rnodes && hz->g.dim == 2)
      hipLaunchKernelGGL((k_dots_hz<NV, TV, U, 2, true>), dim3(w.nrhs, w.nchunk), dim3(256), rbytes, s, w, hz->g,
                         V, hz->z, hz->d, j, rnodes - 1);
    else if (rnodes)
      hipLaunchKernelGGL((k_dots_hz<NV, TV, U, 3, true>), dim3(w.nrhs, w.nchunk), dim3(256), rbytes, s, w, hz->g,
                         V, hz->z, hz->d, j, rnodes - 1);
    else if (hz->g.dim == 2)
      hipLaunchKernelGGL((k_dots_hz<NV, TV, U, 2, false>), dim3(w.nrhs, w.nchunk), dim3(256), 0, s, w, hz->g, V,
                         hz->z, hz->d, j, 0);
    else
      hipLaunchKernelGGL((k_dots_hz<NV, TV, U, 3, false>), dim3(w.nrhs, w.nchunk), dim3(256), 0, s, w, hz->g, V,
                         hz->z, hz->d, j, 0);
  } else if constexpr (std::is_same<TV, double>::value) {
    if (w.lazy)
      hipLaunchKernelGGL((k_dots<NV, true, double>), cgrid(w), dim3(256), 0, s, w, V, wv, j, nullptr);
    else
      hipLaunchKernelGGL((k_dots<NV, false, double>), cgrid(w), dim3(256), 0, s, w, V, wv, j, nullptr);
  }
  if (t1) kt->end(s, hz ? (sv * (j + 1) + 8.0) * Nf * active + 16.0 * Nf : 16.0 * Nf * (j + 2) * active);
  hipLaunchKernelGGL(k_reduce_dots, dim3(j + (w.lazy ? 2 : 1), w.nrhs), dim3(64), 0, s, w, w.h1, j);
  if (t2) kt->begin(s);
  if (c3) {
    if (m3 == 4)
      hipLaunchKernelGGL((k_update_lazy_hz3c<NV, TV, 4>), dim3(w.nrhs, w.nchunk), dim3(256), c3bytes, s, w, hz->g,
                         V, hz->z, hz->d, vnext, j);
    else
      hipLaunchKernelGGL((k_update_lazy_hz3c<NV, TV, 8>), dim3(w.nrhs, w.nchunk), dim3(256), c3bytes, s, w, hz->g,
                         V, hz->z, hz->d, vnext, j);
  } else if (hz) {
    if (rnodes && hz->g.dim == 2)
      hipLaunchKernelGGL((k_update_lazy_hz<NV, TV, U, 2, true>), dim3(w.nrhs, w.nchunk), dim3(256), rbytes, s, w,
                         hz->g, V, hz->z, hz->d, vnext, j, rnodes - 1);
    else if (rnodes)
      hipLaunchKernelGGL((k_update_lazy_hz<NV, TV, U, 3, true>), dim3(w.nrhs, w.nchunk), dim3(256), rbytes, s, w,
                         hz->g, V, hz->z, hz->d, vnext, j, rnodes - 1);
    else if (hz->g.dim == 2)
      hipLaunchKernelGGL((k_update_lazy_hz<NV, TV, U, 2, false>), dim3(w.nrhs, w.nchunk), dim3(256), 0, s, w,
                         hz->g, V, hz->z, hz->d, vnext, j, 0);
    else
      hipLaunchKernelGGL((k_update_lazy_hz<NV, TV, U, 3, false>), dim3(w.nrhs, w.nchunk), dim3(256), 0, s, w,
                         hz->g, V, hz->z, hz->d, vnext, j, 0);
  } else if constexpr (std::is_same<TV, double>::value) {
    if (w.lazy) {
      // w' goes straight into the next basis slot (kept unnormalised)
      hipLaunchKernelGGL((k_update_lazy<NV>), cgrid(w), dim3(256), 0, s, w, V, wv, vnext, j);
    } else {
      hipLaunchKernelGGL((k_update_dots<NV>), cgrid(w), dim3(256), 0, s, w, V, wv, wv, j);
    }
  }
  const double st8 = w.stage ? 8.0 : 0.0;  // staged complex64 copy of w'
  if (t2)
    kt->end(s, hz ? (sv * (j + 2) + 8.0 + st8) * Nf * active + 16.0 * Nf : (16.0 * (j + 3) + st8) * Nf * active);
  if (w.lazy) {
    // second projection only for RHS whose projection cancelled
    hipLaunchKernelGGL(k_fix_flags, dim3(w.nrhs), dim3(64), 0, s, w);
    hipLaunchKernelGGL((k_dots<NV, false, TV>), cgrid(w), dim3(256), 0, s, w, V, (const cx<TV>*)vnext, j, w.fix);
  }
  hipMemcpyAsync(skip_copy, w.done, sizeof(int) * w.nrhs, hipMemcpyDeviceToDevice, s);
  hipLaunchKernelGGL(k_hess_givens, dim3(w.nrhs), dim3(64), 0, s, w, j, tol);
  if constexpr (std::is_same<TV, double>::value) {
    hipLaunchKernelGGL((k_final<NV, double>), cgrid(w), dim3(256), 0, s, w, V,
                       w.lazy ? (const cx<double>*)vnext : wv, vnext, j, skip_copy);
    if (w.lazy && w.stage) hipLaunchKernelGGL(k_restage, cgrid(w), dim3(256), 0, s, w, vnext, skip_copy);
  } else {
    // the corrected vector is stored in complex64 in place: it is the next
    // preconditioner input as it is (no staged copy)
    hipLaunchKernelGGL((k_final<NV, TV>), cgrid(w), dim3(256), 0, s, w, V, (const cx<TV>*)vnext, vnext, j,
                       skip_copy);
  }
}

template <typename TV>
void kr_arnoldi_impl(const KrylovWork& w, const cx<TV>* V, cx<double>* wv, int j, double tol, int* skip_copy,
                     KTimer* kt, int active, hipStream_t s, const HzArgs* hz) {
  if (j < 4)
    arnoldi_nv<4, TV>(w, V, wv, j, tol, skip_copy, kt, active, hz, s);
  else if (j < 6)
    arnoldi_nv<6, TV>(w, V, wv, j, tol, skip_copy, kt, active, hz, s);
  else if (j < 8)
    arnoldi_nv<8, TV>(w, V, wv, j, tol, skip_copy, kt, active, hz, s);
  else if (j < 10)
    arnoldi_nv<10, TV>(w, V, wv, j, tol, skip_copy, kt, active, hz, s);
  else if (j < 12)
    arnoldi_nv<12, TV>(w, V, wv, j, tol, skip_copy, kt, active, hz, s);
  else if (j < 16)
    arnoldi_nv<16, TV>(w, V, wv, j, tol, skip_copy, kt, active, hz, s);
  else
    arnoldi_nv<32, TV>(w, V, wv, j, tol, skip_copy, kt, active, hz, s);
}
template void kr_arnoldi_impl<double>(const KrylovWork&, const cx<double>*, cx<double>*, int, double, int*, KTimer*,
                                      int, hipStream_t, const HzArgs*);
template void kr_arnoldi_impl<float>(const KrylovWork&, const cx<float>*, cx<double>*, int, double, int*, KTimer*,
                                     int, hipStream_t, const HzArgs*);

template <typename TZ>
void kr_update_x(const KrylovWork& w, const cx<TZ>* Z, cx<double>* x, hipStream_t s) {
  hipLaunchKernelGGL(k_backsolve, dim3((w.nrhs + 63) / 64), dim3(64), 0, s, w);
  hipLaunchKernelGGL((k_xupdate<TZ>), cgrid(w), dim3(256), 0, s, w, Z, x);
}
template void kr_update_x<double>(const KrylovWork&, const cx<double>*, cx<double>*, hipStream_t);
template void kr_update_x<float>(const KrylovWork&, const cx<float>*, cx<double>*, hipStream_t);

void kr_true_res(const KrylovWork& w, const cx<double>* r, double* res, double* n2tmp, hipStream_t s) {
  kr_norm2(w, r, n2tmp, nullptr, s);
  hipLaunchKernelGGL(k_div_bnorm, dim3((w.nrhs + 63) / 64), dim3(64), 0, s, w, n2tmp, res);
}

// ------------------------------------------------------------ BiCGSTAB
// Batched right-preconditioned BiCGSTAB (van der Vorst), one independent
// iteration per RHS advanced by the same launches.  Per-RHS scalars live in
// sc[k * nrhs + b]: k = 0 rho, 1 alpha, 2 omega, 3 beta.

// part[c][k][b] = sum over chunk c of conj(a_k) b_k for k < nd (nd <= 2)
__global__ __launch_bounds__(256) void k_bdots(KrylovWork kw, const cx<double>* __restrict__ a0,
                                               const cx<double>* __restrict__ b0, const cx<double>* __restrict__ a1,
                                               const cx<double>* __restrict__ b1, int nd) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (kw.done[b]) return;
  __shared__ double sh[4];
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t base = (int64_t)b * kw.N;
  cx<double> s0 = mk<double>(0.0, 0.0), s1 = mk<double>(0.0, 0.0);
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    s0 = s0 + cmulc(a0[base + p], b0[base + p]);
    if (nd > 1) s1 = s1 + cmulc(a1[base + p], b1[base + p]);
  }
  for (int k = 0; k < nd; ++k) {
    const cx<double> v = k == 0 ? s0 : s1;
    const double re = block_sum(v.re, sh);
    const double im = block_sum(v.im, sh);
    if (threadIdx.x == 0) kw.part[((int64_t)c * (kw.kmax + 1) + k) * kw.nrhs + b] = mk<double>(re, im);
  }
}

__device__ __forceinline__ cx<double> reduce_part(const KrylovWork& kw, int k, int b) {
  double re = 0.0, im = 0.0;
  for (int c = threadIdx.x; c < kw.nchunk; c += 64) {
    const cx<double> v = kw.part[((int64_t)c * (kw.kmax + 1) + k) * kw.nrhs + b];
    re += v.re;
    im += v.im;
  }
  return mk<double>(wave_sum(re), wave_sum(im));
}

__device__ __forceinline__ cx<double> cdiv(cx<double> a, cx<double> b) { return a * cinv(b); }

// scalar stages (one wavefront per RHS):
//   0: rho' = <rhat, r>; beta = (rho'/rho)(alpha/omega) (0 on the first step)
//   1: alpha = rho / <rhat, v>
//   2: omega = <t, s> / <t, t>
//   3: ||r|| estimate, step count, done flag
// A zero denominator (breakdown) ends the RHS's iteration.
__global__ __launch_bounds__(64) void k_bicg_scalar(KrylovWork kw, cx<double>* __restrict__ sc, int stage, int first,
                                                    double tol, int limit) {
  const int b = blockIdx.x;
  if (kw.done[b]) return;
  const int n = kw.nrhs;
  if (stage == 3) {
    double nn = 0.0;
    for (int c = threadIdx.x; c < kw.nchunk; c += 64) nn += kw.partn[(int64_t)c * n + b];
    nn = wave_sum(nn);
    if (threadIdx.x != 0) return;
    const double est = sqrt(nn) / kw.bnorm[b];
    kw.est[b] = est;
    kw.kact[b] += 1;
    if (!(est > tol) || kw.kact[b] >= limit) kw.done[b] = 1;
    return;
  }
  const cx<double> d0 = reduce_part(kw, 0, b);
  const cx<double> d1 = stage == 2 ? reduce_part(kw, 1, b) : mk<double>(0.0, 0.0);
  if (threadIdx.x != 0) return;
  if (stage == 0) {
    if (!(cabs2(d0) > 0.0)) {
      kw.done[b] = 1;
      return;
    }
    const cx<double> beta = first ? mk<double>(0.0, 0.0)
                                  : cdiv(d0, sc[0 * n + b]) * cdiv(sc[1 * n + b], sc[2 * n + b]);
    sc[0 * n + b] = d0;
    sc[3 * n + b] = beta;
  } else if (stage == 1) {
    if (!(cabs2(d0) > 0.0)) {
      kw.done[b] = 1;
      return;
    }
    sc[1 * n + b] = cdiv(sc[0 * n + b], d0);
  } else {
    sc[2 * n + b] = d1.re > 0.0 ? mk<double>(d0.re / d1.re, d0.im / d1.re) : mk<double>(0.0, 0.0);
  }
}

// p = r (first step) or r + beta (p - omega v)
__global__ __launch_bounds__(256) void k_bicg_p(KrylovWork kw, const cx<double>* __restrict__ sc,
                                                const cx<double>* __restrict__ r, cx<double>* __restrict__ p,
                                                const cx<double>* __restrict__ v, int first) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (kw.done[b]) return;
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t base = (int64_t)b * kw.N;
  const int n = kw.nrhs;
  const cx<double> beta = sc[3 * n + b], om = sc[2 * n + b];
  for (int64_t q = p0 + threadIdx.x; q < p1; q += 256) {
    const int64_t o = base + q;
    p[o] = first ? r[o] : r[o] + beta * (p[o] - om * v[o]);
  }
}

// s = r - alpha v
__global__ __launch_bounds__(256) void k_bicg_s(KrylovWork kw, const cx<double>* __restrict__ sc,
                                                const cx<double>* __restrict__ r, const cx<double>* __restrict__ v,
                                                cx<double>* __restrict__ s) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (kw.done[b]) return;
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t base = (int64_t)b * kw.N;
  const cx<double> al = sc[1 * kw.nrhs + b];
  for (int64_t q = p0 + threadIdx.x; q < p1; q += 256) s[base + q] = r[base + q] - al * v[base + q];
}

// x += alpha p^ + omega s^;  r = s - omega t;  partn = ||r||^2
template <typename TZ>
__global__ __launch_bounds__(256) void k_bicg_xr(KrylovWork kw, const cx<double>* __restrict__ sc,
                                                 const cx<TZ>* __restrict__ ph, const cx<TZ>* __restrict__ shat,
                                                 const cx<double>* __restrict__ s, const cx<double>* __restrict__ t,
                                                 cx<double>* __restrict__ x, cx<double>* __restrict__ r) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (kw.done[b]) return;
  __shared__ double sh[4];
  int64_t p0, p1;
  chunk_range(kw.N, kw.nchunk, c, &p0, &p1);
  const int64_t base = (int64_t)b * kw.N;
  const int n = kw.nrhs;
  const cx<double> al = sc[1 * n + b], om = sc[2 * n + b];
  double nn = 0.0;
  for (int64_t q = p0 + threadIdx.x; q < p1; q += 256) {
    const int64_t o = base + q;
    const cx<double> pv = mk<double>((double)ph[o].re, (double)ph[o].im);
    const cx<double> sv = mk<double>((double)shat[o].re, (double)shat[o].im);
    x[o] = x[o] + al * pv + om * sv;
    const cx<double> rv = s[o] - om * t[o];
    r[o] = rv;
    nn += cabs2(rv);
  }
  const double tt = block_sum(nn, sh);
  if (threadIdx.x == 0) kw.partn[(int64_t)c * n + b] = tt;
}

// cycle start: done = conv, step counts cleared
__global__ __launch_bounds__(64) void k_bicg_start(KrylovWork kw, const int* __restrict__ conv) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= kw.nrhs) return;
  kw.done[b] = conv[b];
  kw.kact[b] = 0;
}

void bicg_start(const KrylovWork& w, const int* conv, hipStream_t s) {
  hipLaunchKernelGGL(k_bicg_start, dim3((w.nrhs + 63) / 64), dim3(64), 0, s, w, conv);
}
void bicg_dots(const KrylovWork& w, const cx<double>* a0, const cx<double>* b0, const cx<double>* a1,
               const cx<double>* b1, int nd, hipStream_t s) {
  hipLaunchKernelGGL(k_bdots, cgrid(w), dim3(256), 0, s, w, a0, b0, a1, b1, nd);
}
void bicg_scalar(const KrylovWork& w, cx<double>* sc, int stage, int first, double tol, int limit, hipStream_t s) {
  hipLaunchKernelGGL(k_bicg_scalar, dim3(w.nrhs), dim3(64), 0, s, w, sc, stage, first, tol, limit);
}
void bicg_p(const KrylovWork& w, const cx<double>* sc, const cx<double>* r, cx<double>* p, const cx<double>* v,
            int first, hipStream_t s) {
  hipLaunchKernelGGL(k_bicg_p, cgrid(w), dim3(256), 0, s, w, sc, r, p, v, first);
}
void bicg_s(const KrylovWork& w, const cx<double>* sc, const cx<double>* r, const cx<double>* v, cx<double>* sv,
            hipStream_t s) {
  hipLaunchKernelGGL(k_bicg_s, cgrid(w), dim3(256), 0, s, w, sc, r, v, sv);
}
template <typename TZ>
void bicg_xr(const KrylovWork& w, const cx<double>* sc, const cx<TZ>* ph, const cx<TZ>* shat, const cx<double>* sv,
             const cx<double>* t, cx<double>* x, cx<double>* r, hipStream_t s) {
  hipLaunchKernelGGL((k_bicg_xr<TZ>), cgrid(w), dim3(256), 0, s, w, sc, ph, shat, sv, t, x, r);
}
template void bicg_xr<float>(const KrylovWork&, const cx<double>*, const cx<float>*, const cx<float>*,
                             const cx<double>*, const cx<double>*, cx<double>*, cx<double>*, hipStream_t);
template void bicg_xr<double>(const KrylovWork&, const cx<double>*, const cx<double>*, const cx<double>*,
                              const cx<double>*, const cx<double>*, cx<double>*, cx<double>*, hipStream_t);

// ------------------------------------------------------------ adjoint solves
// With mirror ghosts the Laplacian is L = W^-1 S (S symmetric, W the nodal
// trapezoid weights: 1/2 per boundary axis), so H^H = W conj(H) W^-1 and
// H^H x = b  <=>  x = W conj(y),  H y = W^-1 conj(b).  With BC = 1, W = I.
__device__ __forceinline__ double trap_weight(const Grid& g, int64_t p) {
  if (g.bc != 2) return 1.0;
  const int i = (int)(p % g.n1);
  const int64_t q = p / g.n1;
  const int j = (int)(q % g.n2);
  const int k = (int)(q / g.n2);
  double w = (i == 0 || i == g.n1 - 1) ? 0.5 : 1.0;
  w *= (j == 0 || j == g.n2 - 1) ? 0.5 : 1.0;
  if (g.dim == 3) w *= (k == 0 || k == g.n3 - 1) ? 0.5 : 1.0;
  return w;
}

// mode 0: v = W^-1 conj(v);  mode 1: v = W conj(v);  mode 2: v = W v
__global__ __launch_bounds__(256) void k_adjoint_map(Grid g, cx<double>* __restrict__ v, int nrhs, int mode) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.N) return;
  const double w = trap_weight(g, p);
  const double f = mode == 0 ? 1.0 / w : w;
  for (int b = 0; b < nrhs; ++b) {
    const cx<double> a = v[(int64_t)b * g.N + p];
    v[(int64_t)b * g.N + p] = mk<double>(f * a.re, mode == 2 ? f * a.im : -f * a.im);
  }
}

void adjoint_map(const Grid& g, cx<double>* v, int nrhs, int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_adjoint_map, dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, s, g, v, nrhs, mode);
}

}  // namespace hh
