// GMRES-accelerated damped-Jacobi smoothing ("Jac-GMRES", a relaxType of
// Multigrid.jl that test/ShiftedLaplacianTest.jl:137-142 runs with the
// K-cycle), batched over right-hand sides: s steps of GMRES on A x = b from
// the current x, right-preconditioned by the level's damped-Jacobi factor F:
//
//   r0 = b - A x, beta = ||r0||, v0 = r0 / beta
//   step i:  w = A (F v_i);  classical Gram-Schmidt twice against v_0..v_i
//            (h = V^H w, w -= V h, twice), h_{i+1,i} = ||w||, v_{i+1} = w / ||w||
//   y = argmin || beta e1 - Hbar y ||  (Givens QR per RHS)
//   x <- x + F (V y)
//
// so x minimises ||b - A x|| over x0 + F K_s(A F, r0).  Vectors stay in the
// level's precision; inner products, the Hessenberg matrix and y are
// complex128, and every reduction runs in a fixed order.
#include "hh_kernels.h"

#include <algorithm>

namespace hh {

namespace {

constexpr int GM_NB = GM_MAX_STEPS + 1;  // basis vectors
constexpr int GM_PK = GM_NB + 1;         // partial-sum slots per chunk: dots, then ||w||^2

__device__ __forceinline__ double gsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__device__ __forceinline__ double gbsum(double v, double* sh) {
  v = gsum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  return threadIdx.x == 0 ? sh[0] + sh[1] + sh[2] + sh[3] : 0.0;
}

template <typename T>
__device__ __forceinline__ cx<double> up(cx<T> v) {
  return mk<double>((double)v.re, (double)v.im);
}

int gm_chunks(int64_t N) { return (int)std::max<int64_t>(1, std::min<int64_t>(64, (N + 4095) / 4096)); }

}  // namespace

// part[(c * GM_PK + k) * nrhs + b] = sum over chunk c of conj(V_k) w (k < nv);
// slot GM_NB holds ||w||^2.  V_k is block k of V (stride N * nrhs).
template <typename T>
__global__ __launch_bounds__(256) void k_gm_dots(int64_t N, int nrhs, int nch, const cx<T>* __restrict__ V, int nv,
                                                 const cx<T>* __restrict__ w, cx<double>* __restrict__ part,
                                                 const int* __restrict__ skip) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (skip && skip[b]) return;
  __shared__ double sh[4];
  const int64_t L = (N + nch - 1) / nch, p0 = (int64_t)c * L, p1 = min(N, p0 + L);
  const int64_t base = (int64_t)b * N, stride = N * nrhs;
  cx<double> acc[GM_NB];
#pragma unroll
  for (int k = 0; k < GM_NB; ++k) acc[k] = mk<double>(0.0, 0.0);
  double nn = 0.0;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    const cx<double> wv = up(w[base + p]);
    nn += cabs2(wv);
#pragma unroll
    for (int k = 0; k < GM_NB; ++k)
      if (k < nv) acc[k] = acc[k] + cmulc(up(V[k * stride + base + p]), wv);
  }
#pragma unroll
  for (int k = 0; k < GM_NB; ++k) {
    if (k < nv) {
      const double re = gbsum(acc[k].re, sh), im = gbsum(acc[k].im, sh);
      if (threadIdx.x == 0) part[((int64_t)c * GM_PK + k) * nrhs + b] = mk<double>(re, im);
    }
  }
  const double t = gbsum(nn, sh);
  if (threadIdx.x == 0) part[((int64_t)c * GM_PK + GM_NB) * nrhs + b] = mk<double>(t, 0.0);
}

// Reductions over chunks, one wavefront per RHS.  State per RHS b:
//   hc[k * nrhs + b]  projection coefficients of the current pass
//   H[b * GM_NB * GM_MAX_STEPS + k * GM_MAX_STEPS + i]  Hessenberg (row k, column i)
//   sc[b]  1 / norm for the next scaling (0 when the norm is 0)
//   beta[b]  ||r0||
// mode 0: beta = ||r0||, sc = 1 / beta
// mode 1: first projection: hc = V^H w (k < nv)
// mode 2: second projection: hc = V^H w, H[:, col] = first + second
//         (the first pass's coefficients are in H[:, col] already)
// mode 3: norm: H[col + 1, col] = ||w||, sc = 1 / ||w||
__global__ __launch_bounds__(64) void k_gm_reduce(int nrhs, int nch, const cx<double>* __restrict__ part, int mode,
                                                  int nv, int col, cx<double>* __restrict__ hc,
                                                  cx<double>* __restrict__ H, double* __restrict__ sc,
                                                  double* __restrict__ beta, const int* __restrict__ skip) {
  const int b = blockIdx.x;
  if (skip && skip[b]) return;
  cx<double>* Hb = H + (int64_t)b * GM_NB * GM_MAX_STEPS;
  if (mode == 0 || mode == 3) {
    double nn = 0.0;
    for (int c = threadIdx.x; c < nch; c += 64) nn += part[((int64_t)c * GM_PK + GM_NB) * nrhs + b].re;
    nn = gsum(nn);
    if (threadIdx.x != 0) return;
    const double nr = nn > 0.0 ? sqrt(nn) : 0.0;
    sc[b] = nr > 0.0 ? 1.0 / nr : 0.0;
    if (mode == 0)
      beta[b] = nr;
    else
      Hb[(col + 1) * GM_MAX_STEPS + col] = mk<double>(nr, 0.0);
    return;
  }
  for (int k = 0; k < nv; ++k) {
    double re = 0.0, im = 0.0;
    for (int c = threadIdx.x; c < nch; c += 64) {
      const cx<double> v = part[((int64_t)c * GM_PK + k) * nrhs + b];
      re += v.re;
      im += v.im;
    }
    re = gsum(re);
    im = gsum(im);
    if (threadIdx.x == 0) {
      const cx<double> h = mk<double>(re, im);
      hc[k * nrhs + b] = h;
      cx<double>& e = Hb[k * GM_MAX_STEPS + col];
      e = mode == 1 ? h : e + h;
    }
  }
}

// w -= sum_k hc_k V_k (k < nv)
template <typename T>
__global__ __launch_bounds__(256) void k_gm_update(int64_t N, int nrhs, const cx<T>* __restrict__ V, int nv,
                                                   const cx<double>* __restrict__ hc, cx<T>* __restrict__ w,
                                                   const int* __restrict__ skip) {
  const int b = blockIdx.y;
  if (skip && skip[b]) return;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  const int64_t o = (int64_t)b * N + p, stride = N * nrhs;
  cx<double> v = up(w[o]);
  for (int k = 0; k < nv; ++k) v = v - hc[k * nrhs + b] * up(V[k * stride + o]);
  w[o] = mk<T>((T)v.re, (T)v.im);
}

// y = sc[b] x (per RHS), in place allowed
template <typename T>
__global__ __launch_bounds__(256) void k_gm_scale(int64_t N, int nrhs, const double* __restrict__ sc,
                                                  const cx<T>* x, cx<T>* y, const int* __restrict__ skip) {
  const int b = blockIdx.y;
  if (skip && skip[b]) return;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  const int64_t o = (int64_t)b * N + p;
  const double s = sc[b];
  const cx<T> v = x[o];
  y[o] = mk<T>((T)(s * (double)v.re), (T)(s * (double)v.im));
}

// u = F o v (the node's damped-Jacobi factor times v)
template <typename T>
__global__ __launch_bounds__(256) void k_gm_hadamard(int64_t N, int nrhs, const cx<T>* __restrict__ f,
                                                     const cx<T>* __restrict__ v, cx<T>* __restrict__ u,
                                                     const int* __restrict__ skip) {
  const int b = blockIdx.y;
  if (skip && skip[b]) return;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  const int64_t o = (int64_t)b * N + p;
  u[o] = f[p] * v[o];
}

// y = argmin || beta e1 - Hbar y || over the leading columns whose
// subdiagonal stays nonzero (an exact Krylov direction ends the space), by
// Givens rotations; one thread per RHS.
__global__ __launch_bounds__(64) void k_gm_lsq(int nrhs, int s, const cx<double>* __restrict__ H,
                                               const double* __restrict__ beta, cx<double>* __restrict__ y,
                                               const int* __restrict__ skip) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= nrhs || (skip && skip[b])) return;
  cx<double> R[GM_NB][GM_MAX_STEPS], g[GM_NB];
  const cx<double>* Hb = H + (int64_t)b * GM_NB * GM_MAX_STEPS;
  for (int k = 0; k <= s; ++k) {
    g[k] = mk<double>(0.0, 0.0);
    for (int i = 0; i < s; ++i) R[k][i] = Hb[k * GM_MAX_STEPS + i];
  }
  g[0] = mk<double>(beta[b], 0.0);
  int kk = s;  // columns in use
  for (int i = 0; i < s; ++i)
    if (!(cabs2(R[i + 1][i]) > 0.0)) {
      kk = i + 1;
      break;
    }
  cx<double>* yb = y + (int64_t)b * GM_MAX_STEPS;
  for (int i = 0; i < GM_MAX_STEPS; ++i) yb[i] = mk<double>(0.0, 0.0);
  if (!(beta[b] > 0.0)) return;
  for (int i = 0; i < kk; ++i) {
    // rotation zeroing R[i+1][i] against R[i][i]
    const cx<double> a = R[i][i], c = R[i + 1][i];
    const double na = sqrt(cabs2(a)), nc = sqrt(cabs2(c));
    if (!(nc > 0.0)) continue;
    const double r = sqrt(na * na + nc * nc);
    // [cs  sn; -conj(sn) cs] with cs real: cs = |a| / r, sn = (a / |a|) conj(c) / r
    const double cs = na / r;
    const cx<double> ph = na > 0.0 ? mk<double>(a.re / na, a.im / na) : mk<double>(1.0, 0.0);
    const cx<double> sn = ph * mk<double>(c.re / r, -c.im / r);
    for (int j = i; j < kk; ++j) {
      const cx<double> t0 = R[i][j], t1 = R[i + 1][j];
      R[i][j] = cs * t0 + sn * t1;
      R[i + 1][j] = cs * t1 - mk<double>(sn.re, -sn.im) * t0;
    }
    const cx<double> g0 = g[i], g1 = g[i + 1];
    g[i] = cs * g0 + sn * g1;
    g[i + 1] = cs * g1 - mk<double>(sn.re, -sn.im) * g0;
  }
  for (int i = kk - 1; i >= 0; --i) {
    cx<double> t = g[i];
    for (int j = i + 1; j < kk; ++j) t = t - R[i][j] * yb[j];
    yb[i] = cabs2(R[i][i]) > 0.0 ? t * cinv(R[i][i]) : mk<double>(0.0, 0.0);
  }
}

// x (+)= F o (sum_i y_i V_i) over the first s basis vectors; assign = 1 writes
// instead of adding (smoothing from x = 0)
template <typename T>
__global__ __launch_bounds__(256) void k_gm_xupdate(int64_t N, int nrhs, const cx<T>* __restrict__ V, int s,
                                                    const cx<double>* __restrict__ y, const cx<T>* __restrict__ f,
                                                    cx<T>* __restrict__ x, int assign, const int* __restrict__ skip) {
  const int b = blockIdx.y;
  if (skip && skip[b]) return;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  const int64_t o = (int64_t)b * N + p, stride = N * nrhs;
  const cx<double>* yb = y + (int64_t)b * GM_MAX_STEPS;
  cx<double> acc = mk<double>(0.0, 0.0);
  for (int i = 0; i < s; ++i) acc = acc + yb[i] * up(V[i * stride + o]);
  const cx<double> e = up(f[p]) * acc;
  const cx<double> xn = assign ? e : up(x[o]) + e;
  x[o] = mk<T>((T)xn.re, (T)xn.im);
}

size_t gm_part_elems(int nrhs) { return (size_t)64 * GM_PK * nrhs; }

template <typename T>
void gm_smooth(const GmArgs<T>& a, int steps, hipStream_t st) {
  if (steps <= 0) return;
  const int64_t N = a.g.N;
  const int nrhs = a.nrhs, nch = gm_chunks(N);
  const int64_t stride = N * nrhs;
  const dim3 eg((unsigned)((N + 255) / 256), nrhs), cg(nch, nrhs);
  cx<T>* V = a.V;
  // r0 = b - A x (or b from x = 0), beta, v0 = r0 / beta
  if (a.zero_x)
    (void)hipMemcpyAsync(V, a.b, sizeof(cx<T>) * stride, hipMemcpyDeviceToDevice, st);  // checked by the caller
  else
    launch_stencil<T, T, T, M_RESID>(a.g, nrhs, a.x, a.b, V, a.d, nullptr, mk<double>(0, 0), 0.0, a.skip, st);
  hipLaunchKernelGGL((k_gm_dots<T>), cg, dim3(256), 0, st, N, nrhs, nch, V, 0, V, a.part, a.skip);
  hipLaunchKernelGGL(k_gm_reduce, dim3(nrhs), dim3(64), 0, st, nrhs, nch, a.part, 0, 0, 0, a.hc, a.H, a.sc, a.beta,
                     a.skip);
  hipLaunchKernelGGL((k_gm_scale<T>), eg, dim3(256), 0, st, N, nrhs, a.sc, V, V, a.skip);
  for (int i = 0; i < steps; ++i) {
    cx<T>* vi = V + i * stride;
    cx<T>* w = V + (i + 1) * stride;
    hipLaunchKernelGGL((k_gm_hadamard<T>), eg, dim3(256), 0, st, N, nrhs, a.f, vi, a.u, a.skip);
    launch_stencil<T, T, T, M_APPLY>(a.g, nrhs, a.u, nullptr, w, a.d, nullptr, mk<double>(0, 0), 0.0, a.skip, st);
    for (int pass = 1; pass <= 2; ++pass) {
      hipLaunchKernelGGL((k_gm_dots<T>), cg, dim3(256), 0, st, N, nrhs, nch, V, i + 1, w, a.part, a.skip);
      hipLaunchKernelGGL(k_gm_reduce, dim3(nrhs), dim3(64), 0, st, nrhs, nch, a.part, pass, i + 1, i, a.hc, a.H,
                         a.sc, a.beta, a.skip);
      hipLaunchKernelGGL((k_gm_update<T>), eg, dim3(256), 0, st, N, nrhs, V, i + 1, a.hc, w, a.skip);
    }
    hipLaunchKernelGGL((k_gm_dots<T>), cg, dim3(256), 0, st, N, nrhs, nch, V, 0, w, a.part, a.skip);
    hipLaunchKernelGGL(k_gm_reduce, dim3(nrhs), dim3(64), 0, st, nrhs, nch, a.part, 3, 0, i, a.hc, a.H, a.sc, a.beta,
                       a.skip);
    if (i + 1 < steps) hipLaunchKernelGGL((k_gm_scale<T>), eg, dim3(256), 0, st, N, nrhs, a.sc, w, w, a.skip);
  }
  hipLaunchKernelGGL(k_gm_lsq, dim3((nrhs + 63) / 64), dim3(64), 0, st, nrhs, steps, a.H, a.beta, a.y, a.skip);
  hipLaunchKernelGGL((k_gm_xupdate<T>), eg, dim3(256), 0, st, N, nrhs, V, steps, a.y, a.f, a.x, a.zero_x ? 1 : 0,
                     a.skip);
}

template void gm_smooth<float>(const GmArgs<float>&, int, hipStream_t);
template void gm_smooth<double>(const GmArgs<double>&, int, hipStream_t);

}  // namespace hh
