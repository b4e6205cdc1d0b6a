// Fused multigrid legs for 3-D levels with mirror ghosts (BC = 2), complex64.
//
// Down leg (pre-smoothing from zero, then the restricted residual):
//
//   x1 = F b,  x2 = x1 + F (b - A x1),  b_c = R (b - A x2)
//
// reads b, d, F once and writes x2 and b_c once (unfused: the zero-start
// sweep, one sweep, the residual and the full weighting, four passes).  Up
// leg (last coarse correction, then post-smoothing):
//
//   x' = x + P e_c, then two damped-Jacobi sweeps x <- x + F (b - A x)
//
// reads x, b, F and the coarse correction once and writes the result once
// (unfused: the prolongation and two sweeps, three passes).  On W levels the
// middle step between the two coarse corrections,
//
//   x' = x + P e_c,  b_c = R (b - A x')
//
// is one pass too (unfused: the prolongation, the residual and the full
// weighting).  F = relax D^-1
// with D the operator's diagonal, so a sweep is (1 - relax) x + F (b + ih1 sx
// + ih2 sy + ih3 sz) with s the neighbour sums along each axis and reads no
// diagonal; the residual keeps the operator's form b - A x (a split residual
// cancels terms of size 6 |x| / h^2 down to the residual).
//
// A block owns a 28 x 28 column of nodes (a 32 x 32 in-plane tile with a
// 2-node halo, one position per thread of 1024: a wavefront covers two
// 32-node tile rows) and marches it along z over a run of output planes.  (A
// 64-wide tile wastes a quarter of its lanes on 193-wide planes; 512 threads
// with two positions each need 92-98 VGPRs, 16 resident wavefronts per CU
// against 32 at 55-60.)  Step z forms the first stage on plane z, the second
// on plane z - 1 and the third on plane z - 2; the earlier stages' values of
// the last two planes live in LDS rings (in-plane neighbours), a thread's
// own z-neighbours in registers, and one barrier per step separates every
// LDS write from its readers.  A run starting at z0 > 0 first forms its
// earlier stages on the planes before z0 (recomputed, not exchanged), so
// runs are independent.  Ghost planes
// and in-plane ghosts mirror (index -1 = 1, n = n - 2): a ghost node's value
// is its mirror image's at every stage (up to the rounding of its neighbour
// sums, taken in the other order), so every formula is the unfused one.
#include "hh_kernels.h"

namespace hh {

namespace {
constexpr int M3_W = 32, M3_RY = 32, M3_NT = 1024, M3_NP = M3_W * M3_RY;
// owned columns / rows of a tile with an H-node halo (H = 2: two stencil
// stages, or one stage and the full weighting; the down leg's two stages and
// full weighting need H = 3)
constexpr int m3_tx(int H) { return M3_W - 2 * H; }
constexpr int m3_ty(int H) { return M3_RY - 2 * H; }
// NT threads: NT / 32 tile rows per pass, KR rows per thread
template <int NT>
struct Geo3 {
  static constexpr int RPP = NT / M3_W, KR = M3_RY / RPP;
  static_assert(M3_RY % RPP == 0, "tile rows split evenly over the threads");
};

// a block's tile, z-run and this thread's in-plane positions
template <int KR>
struct March3 {
  int rhs, z0, z1, za, zb;  // output planes [z0, z1), first-stage planes [za, zb)
  int fx0, fy0, gi;         // first owned node of the tile; this thread's (mirrored) x index
  int lj[KR], gj[KR];       // tile rows; their (mirrored) y indices
  int po[KR];               // in-plane offsets of the (mirrored) positions (a plane has < 2^31 nodes)
  bool in1[KR];             // inside the halo-1 ring (second stage)
  bool own[KR];             // owned node (third stage, stored)
};

template <int NT, int H, int KR>
__device__ __forceinline__ bool march3_setup(const TileMap& tm, int ytiles, int zrun, const Grid& g,
                                             const int* __restrict__ skip, March3<KR>& m) {
  constexpr int TX = m3_tx(H), TY = m3_ty(H);
  constexpr int RPP = Geo3<NT>::RPP;
  int xt, yz;
  tile_of(tm, xt, yz, m.rhs);
  if (skip && skip[m.rhs]) return false;
  const int yt = yz % ytiles, zc = yz / ytiles;
  m.z0 = zc * zrun;
  m.z1 = min(g.n3, m.z0 + zrun);
  m.za = max(0, m.z0 - H);
  m.zb = min(g.n3, m.z1 + H);
  m.fx0 = xt * TX;
  m.fy0 = yt * TY;
  const int lane = threadIdx.x & (M3_W - 1), row = threadIdx.x / M3_W;
  const int gx = m.fx0 - H + lane;
  m.gi = mirror_idx(gx, g.n1);
  const bool inx = lane >= 1 && lane < M3_W - 1, ownx = lane >= H && lane < H + TX && gx < g.n1;
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int lj = row + RPP * k, gy = m.fy0 - H + lj;
    m.lj[k] = lj;
    m.gj[k] = mirror_idx(gy, g.n2);
    m.po[k] = m.gj[k] * g.n1 + m.gi;
    m.in1[k] = inx && lj >= 1 && lj < M3_RY - 1;
    m.own[k] = ownx && lj >= H && lj < H + TY && gy < g.n2;
  }
  return true;
}

// owned node's offset in its plane (only meaningful where own[k])
template <int H, int KR>
__device__ __forceinline__ int64_t own_off(const March3<KR>& m, int k, int n1) {
  return (int64_t)(m.fy0 - H + m.lj[k]) * n1 + (m.fx0 - H + (int)(threadIdx.x & (M3_W - 1)));
}

// bilinear interpolant of coarse plane K at the fine in-plane position
// (gi, gj): the coarse nodes around it, weights 1 or 1/2 per axis
template <typename T>
__device__ __forceinline__ cx<T> coarse_at(const cx<T>* __restrict__ cr, const Grid& gc, int K, int gi, int gj) {
  const cx<T>* cp = cr + (int64_t)K * gc.n1 * gc.n2;
  const int ci0 = gi >> 1, ci1 = (gi & 1) ? ci0 + 1 : ci0;
  const int cj0 = gj >> 1, cj1 = (gj & 1) ? cj0 + 1 : cj0;
  const T w = ((gi & 1) ? (T)0.5 : (T)1) * ((gj & 1) ? (T)0.5 : (T)1);
  const int64_t r0 = (int64_t)cj0 * gc.n1, r1 = (int64_t)cj1 * gc.n1;
  cx<T> acc = w * cp[r0 + ci0];
  if (gi & 1) acc = acc + w * cp[r0 + ci1];
  if (gj & 1) {
    acc = acc + w * cp[r1 + ci0];
    if (gi & 1) acc = acc + w * cp[r1 + ci1];
  }
  return acc;
}

// damped-Jacobi sweep at one node without the diagonal:
// (1 - relax) xc + F (b + ih1 (xl + xr) + ih2 (xm + xp) + ih3 (xd + xu))
template <typename T>
__device__ __forceinline__ cx<T> jac7(cx<T> xc, cx<T> xl, cx<T> xr, cx<T> xm, cx<T> xp, cx<T> xd, cx<T> xu,
                                      cx<T> bq, cx<T> fq, T ih1, T ih2, T ih3, T omr) {
  const T tr = fma(ih1, xl.re + xr.re, fma(ih2, xm.re + xp.re, fma(ih3, xd.re + xu.re, bq.re)));
  const T ti = fma(ih1, xl.im + xr.im, fma(ih2, xm.im + xp.im, fma(ih3, xd.im + xu.im, bq.im)));
  return mk<T>(fma(omr, xc.re, fma(fq.re, tr, -(fq.im * ti))), fma(omr, xc.im, fma(fq.re, ti, fq.im * tr)));
}

}  // namespace

// Down leg, with the full weighting of its residual (b_c = R r written, r
// itself never stored).  Step z: x1 on plane z, x2 on plane z - 1 (stored for
// the run's own planes), r on plane z - 2 over the tile less its two outer
// rings (kept in LDS), and the in-plane 9-point sums of plane z - 3's r at
// the tile's coarse nodes, each coarse plane K summed from its fine planes
// 2K - 1, 2K, 2K + 1 in registers and stored, as in the middle step.  The
// weights reach one fine node past the owned range, where r needs a 3-node
// halo: the tile owns 26 x 26 fine nodes (13 x 13 coarse nodes, starting at
// an even corner) and a run recomputes three planes on each side.  (On
// 193 x 193 x 97 that is 31% more tiles than 28 x 28 owned nodes give, and
// still faster than storing r for a separate restriction: 0.96 against
// 1.02 + 0.26 ms.)
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_down_fused3(TileMap tm, int ytiles, int zrun, Grid gf, Grid gc,
                                                      const cx<T>* __restrict__ b, const cx<T>* __restrict__ d,
                                                      const cx<T>* __restrict__ f, cx<T>* __restrict__ xo,
                                                      cx<T>* __restrict__ bco, const int* __restrict__ skip, T omr) {
  constexpr int KR = Geo3<NT>::KR, W = M3_W, H = 3;
  static_assert(KR == 1, "one tile position per thread");
  __shared__ cx<T> ra[2][M3_NP], rs[2][M3_NP], rr[2][M3_NP];
  March3<KR> m;
  if (!march3_setup<NT, H>(tm, ytiles, zrun, gf, skip, m)) return;
  const int n1 = gf.n1, n3 = gf.n3;
  const int64_t plane = (int64_t)n1 * gf.n2, off = (int64_t)m.rhs * gf.N;
  const cx<T>* __restrict__ br = b + off;
  cx<T>* __restrict__ xr = xo + off;
  cx<T>* __restrict__ bc = bco + (int64_t)m.rhs * gc.N;
  const T ih1 = (T)gf.ih1, ih2 = (T)gf.ih2, ih3 = (T)gf.ih3;
  const int lane = threadIdx.x & (W - 1), lj = m.lj[0], qq = lj * W + lane;
  const bool in2 = lane >= 2 && lane < W - 2 && lj >= 2 && lj < M3_RY - 2;  // r's positions
  // this thread's coarse node, if any: fine tile position (H + 2 ci, H + 2 cj)
  const int ci = (lane - H) >> 1, cj = (lj - H) >> 1;
  const int I = (m.fx0 >> 1) + ci, J = (m.fy0 >> 1) + cj;
  const bool cown = lane >= H && lane < H + m3_tx(H) && !((lane - H) & 1) && lj >= H && lj < H + m3_ty(H) &&
                    !((lj - H) & 1) && I < gc.n1 && J < gc.n2;
  const int64_t cplane = (int64_t)gc.n1 * gc.n2, ccol = (int64_t)J * gc.n1 + I;
  const T w[3] = {(T)0.25, (T)0.5, (T)0.25};
  // planes of each stage: x2 on [p0, p1), r on [r0, r1) (the coarse planes of
  // the run need r one plane beyond it on each side)
  const int p0 = max(0, m.z0 - 2), p1 = min(n3, m.z1 + 2), r0 = max(0, m.z0 - 1), r1 = min(n3, m.z1 + 1);
  // b, F of plane z (b0, f0) and z - 1 (b1, f1), b of z - 2 (b2); x1 of
  // planes z, z - 1, z - 2; x2 of planes z - 1, z - 2, z - 3; in-plane sums
  // of r on planes q - 1, q - 2
  cx<T> b0, f0, b1 = mk<T>(0, 0), f1 = b1, b2 = b1, dq;
  cx<T> xa0 = b1, xa1 = b1, xa2 = b1, sa0 = b1, sa1 = b1, sa2 = b1, s1 = b1, s2 = b1;
  cx<T> bn, fn;  // the next step's loads
  auto load_step = [&](int zz) {  // b, F of plane zz (clamped: always in bounds)
    const int64_t pz = (int64_t)min(zz, n3 - 1) * plane;
    bn = br[pz + m.po[0]];
    fn = f[pz + m.po[0]];
  };
  load_step(m.za);
  for (int z = m.za; z <= r1 + 2; ++z) {
    dq = d[(int64_t)min(max(z - 2, 0), n3 - 1) * plane + m.po[0]];  // the diagonal of plane z - 2
    b0 = bn;
    f0 = fn;
    if (z < r1 + 2) load_step(z + 1);
    if (z < m.zb) {  // x1 = F b on plane z
      xa0 = cmul_fma(f0, b0);
      ra[z & 1][qq] = xa0;
    }
    const int p = z - 1;
    if (p >= p0 && p < p1 && m.in1[0]) {  // x2 on plane p
      const cx<T>* a = ra[p & 1];
      const bool lo = p == 0, hi = p == n3 - 1;
      const cx<T> xm = lo ? xa0 : xa2, xp = hi ? xa2 : xa0;
      sa0 = jac7(xa1, a[qq - 1], a[qq + 1], a[qq - W], a[qq + W], xm, xp, b1, f1, ih1, ih2, ih3, omr);
      rs[p & 1][qq] = sa0;
      if (p >= m.z0 && p < m.z1 && m.own[0]) xr[(int64_t)p * plane + own_off<H>(m, 0, n1)] = sa0;
    }
    const int r = z - 2;
    if (r >= r0 && r < r1 && in2) {  // r = b - A x2 on plane r
      const cx<T>* sx = rs[r & 1];
      const bool lo = r == 0, hi = r == n3 - 1;
      const cx<T> sm = lo ? sa0 : sa2, sp = hi ? sa2 : sa0;
      rr[r & 1][qq] = b2 - apply7(sa1, sx[qq - 1], sx[qq + 1], sx[qq - W], sx[qq + W], sm, sp, ih1, ih2, ih3, dq);
    }
    const int q = z - 3;
    if (q >= r0 && q < r1 && cown) {  // full weighting of plane q's r
      const cx<T>* rq = rr[q & 1];
      cx<T> sq = mk<T>(0, 0);
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int aa = 0; aa < 3; ++aa) sq = sq + (w[bb] * w[aa]) * rq[(lj - 1 + bb) * W + (lane - 1 + aa)];
      if (q & 1) {  // coarse plane K = (q - 1) / 2 is complete
        const int K = (q - 1) >> 1;
        if (2 * K >= m.z0) {
          const cx<T> sm = K == 0 ? sq : s2;  // plane -1 mirrors plane 1
          bc[(int64_t)K * cplane + ccol] = w[0] * sm + w[1] * s1 + w[2] * sq;
        }
      } else if (q == n3 - 1) {  // the last coarse plane: plane n3 mirrors n3 - 2
        bc[(int64_t)(q >> 1) * cplane + ccol] = w[0] * s1 + w[1] * sq + w[2] * s1;
      }
      s2 = s1;
      s1 = sq;
    }
    b2 = b1;
    b1 = b0;
    f1 = f0;
    xa2 = xa1;
    xa1 = xa0;
    sa2 = sa1;
    sa1 = sa0;
    __syncthreads();
  }
}

// Up leg.  Step z: x' on plane z, the first sweep on plane z - 1, the second
// on plane z - 2 (owned nodes, run's own planes, stored).  x of plane z and b,
// F of plane z - 1 are loaded a step ahead; b, F stay one more step in
// registers for the second sweep.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_up_fused3(TileMap tm, int ytiles, int zrun, Grid gf, Grid gc,
                                                   const cx<T>* __restrict__ x, const cx<T>* __restrict__ ec,
                                                   const cx<T>* __restrict__ b, const cx<T>* __restrict__ f,
                                                   cx<T>* __restrict__ xo, const int* __restrict__ skip, T omr) {
  constexpr int KR = Geo3<NT>::KR, W = M3_W;
  __shared__ cx<T> ra[2][M3_NP], rs[2][M3_NP];
  March3<KR> m;
  if (!march3_setup<NT, 2>(tm, ytiles, zrun, gf, skip, m)) return;
  const int n1 = gf.n1, n3 = gf.n3;
  const int64_t plane = (int64_t)n1 * gf.n2, off = (int64_t)m.rhs * gf.N;
  const cx<T>* __restrict__ xr = x + off;
  const cx<T>* __restrict__ br = b + off;
  const cx<T>* __restrict__ cr = ec + (int64_t)m.rhs * gc.N;
  cx<T>* __restrict__ yr = xo + off;
  const T ih1 = (T)gf.ih1, ih2 = (T)gf.ih2, ih3 = (T)gf.ih3;
  // in-plane bilinear interpolant of coarse plane K at this thread's positions
  auto cval = [&](int K, int k) { return coarse_at(cr, gc, K, m.gi, m.gj[k]); };
  cx<T> cK[KR];
#pragma unroll
  for (int k = 0; k < KR; ++k) cK[k] = cval(m.za >> 1, k);
  // b, F of plane z - 1 (bp, fp) and z - 2 (b2, f2); x' of planes z, z - 1,
  // z - 2; first-sweep values of planes z - 1, z - 2, z - 3
  cx<T> xv[KR], bp[KR], fp[KR], b2[KR], f2[KR];
  cx<T> xa0[KR], xa1[KR], xa2[KR], sa0[KR], sa1[KR], sa2[KR];
  cx<T> xn[KR], bn[KR], fn[KR];  // the next step's loads
  auto load_step = [&](int zz) {  // x of plane zz, b and F of plane zz - 1 (clamped: always in bounds)
    const int64_t pz = (int64_t)min(zz, n3 - 1) * plane, pb = (int64_t)max(zz - 1, 0) * plane;
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      xn[k] = xr[pz + m.po[k]];
      bn[k] = br[pb + m.po[k]];
      fn[k] = f[pb + m.po[k]];
    }
  };
  load_step(m.za);
  for (int z = m.za; z <= m.z1 + 1; ++z) {
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      xv[k] = xn[k];
      bp[k] = bn[k];
      fp[k] = fn[k];
    }
    if (z < m.z1 + 1) load_step(z + 1);
    if (z < m.zb) {  // x' = x + P e_c on plane z
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        cx<T> e;
        if (z & 1) {
          const cx<T> cN = cval((z >> 1) + 1, k);
          e = (T)0.5 * (cK[k] + cN);
          cK[k] = cN;
        } else {
          e = cK[k];
        }
        xa0[k] = xv[k] + e;
        ra[z & 1][m.lj[k] * W + (threadIdx.x & (W - 1))] = xa0[k];
      }
    }
    const int p = z - 1;
    if (p >= max(0, m.z0 - 1) && p < min(n3, m.z1 + 1)) {  // first sweep on plane p
      const cx<T>* a = ra[p & 1];
      const bool lo = p == 0, hi = p == n3 - 1;
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        if (m.in1[k]) {
          const int qq = m.lj[k] * W + (threadIdx.x & (W - 1));
          const cx<T> xm = lo ? xa0[k] : xa2[k], xp = hi ? xa2[k] : xa0[k];
          sa0[k] = jac7(xa1[k], a[qq - 1], a[qq + 1], a[qq - W], a[qq + W], xm, xp, bp[k], fp[k], ih1, ih2, ih3, omr);
          rs[p & 1][qq] = sa0[k];
        }
      }
    }
    const int r = z - 2;
    if (r >= m.z0 && r < m.z1) {  // second sweep on plane r, owned nodes stored
      const cx<T>* s = rs[r & 1];
      const bool lo = r == 0, hi = r == n3 - 1;
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        if (m.own[k]) {
          const int qq = m.lj[k] * W + (threadIdx.x & (W - 1));
          const cx<T> sm = lo ? sa0[k] : sa2[k], sp = hi ? sa2[k] : sa0[k];
          yr[(int64_t)r * plane + own_off<2>(m, k, n1)] =
              jac7(sa1[k], s[qq - 1], s[qq + 1], s[qq - W], s[qq + W], sm, sp, b2[k], f2[k], ih1, ih2, ih3, omr);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      b2[k] = bp[k];
      f2[k] = fp[k];
      xa2[k] = xa1[k];
      xa1[k] = xa0[k];
      sa2[k] = sa1[k];
      sa1[k] = sa0[k];
    }
    __syncthreads();
  }
}

// Middle step of a W visit (the first coarse correction added, then the
// residual for the second and its full weighting).  Step z: x' = x + P e_c
// on plane z (every tile position; owned nodes of the run's planes stored),
// r = b - A x' on plane z - 1 (the tile less its outer ring, kept in LDS),
// and the in-plane 9-point weighted sum of plane z - 2's r at the tile's
// coarse nodes, each coarse plane K summed from its fine planes 2K - 1, 2K,
// 2K + 1 in registers and stored.  The tile's 28 x 28 owned fine nodes start
// at an even corner, so it owns 14 x 14 coarse nodes whose weights reach one
// fine node below the owned range (inside the 2-node halo); a run of fine
// planes [z0, z1) (z0 even) owns the coarse planes K with 2K in [z0, z1).
// The sums are formed in the order of the separate restriction kernel; at
// the grid's edges the residual one node beyond the edge is the tile's own
// value at that mirror position (the mirror node's value up to rounding, the
// stencil's neighbours being summed in the other order).
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_mid_fused3(TileMap tm, int ytiles, int zrun, Grid gf, Grid gc,
                                                    const cx<T>* __restrict__ x, const cx<T>* __restrict__ ec,
                                                    const cx<T>* __restrict__ b, const cx<T>* __restrict__ d,
                                                    cx<T>* __restrict__ xo, cx<T>* __restrict__ bco,
                                                    const int* __restrict__ skip) {
  constexpr int KR = Geo3<NT>::KR, W = M3_W, H = 2;
  static_assert(KR == 1, "one tile position per thread");
  __shared__ cx<T> ra[2][M3_NP], rr[2][M3_NP];
  March3<KR> m;
  if (!march3_setup<NT, H>(tm, ytiles, zrun, gf, skip, m)) return;
  const int n1 = gf.n1, n3 = gf.n3;
  const int zb = min(n3, m.z1 + 1);  // x' on [za, zb): r on [z0 - 1, z1) needs one plane each side
  const int64_t plane = (int64_t)n1 * gf.n2, off = (int64_t)m.rhs * gf.N;
  const cx<T>* __restrict__ xr = x + off;
  const cx<T>* __restrict__ br = b + off;
  const cx<T>* __restrict__ cr = ec + (int64_t)m.rhs * gc.N;
  cx<T>* __restrict__ yr = xo + off;
  cx<T>* __restrict__ bc = bco + (int64_t)m.rhs * gc.N;
  const T ih1 = (T)gf.ih1, ih2 = (T)gf.ih2, ih3 = (T)gf.ih3;
  const int lane = threadIdx.x & (W - 1), lj = m.lj[0], qq = lj * W + lane;
  // this thread's coarse node, if any: fine tile position (H + 2 ci, H + 2 cj)
  const int ci = (lane - H) >> 1, cj = (lj - H) >> 1;
  const int I = (m.fx0 >> 1) + ci, J = (m.fy0 >> 1) + cj;
  const bool cown = lane >= H && lane < H + m3_tx(H) && !((lane - H) & 1) && lj >= H && lj < H + m3_ty(H) &&
                    !((lj - H) & 1) && I < gc.n1 && J < gc.n2;
  const int64_t cplane = (int64_t)gc.n1 * gc.n2, ccol = (int64_t)J * gc.n1 + I;
  const T w[3] = {(T)0.25, (T)0.5, (T)0.25};
  cx<T> cK = coarse_at(cr, gc, m.za >> 1, m.gi, m.gj[0]);
  // x' of planes z, z - 1, z - 2; in-plane sums of r on planes q - 1, q - 2
  cx<T> xa0 = mk<T>(0, 0), xa1 = xa0, xa2 = xa0, s1 = xa0, s2 = xa0;
  cx<T> xn, bn, dn;  // the next step's loads: x of plane zz, b and d of plane zz - 1
  auto load_step = [&](int zz) {  // clamped plane indices: always in bounds
    const int64_t pz = (int64_t)min(zz, n3 - 1) * plane, pb = (int64_t)max(min(zz, n3) - 1, 0) * plane;
    xn = xr[pz + m.po[0]];
    bn = br[pb + m.po[0]];
    dn = d[pb + m.po[0]];
  };
  load_step(m.za);
  for (int z = m.za; z <= m.z1 + 1; ++z) {
    const cx<T> xv = xn, bp = bn, dp = dn;
    if (z < m.z1 + 1) load_step(z + 1);
    if (z < zb) {  // x' = x + P e_c on plane z
      cx<T> e;
      if (z & 1) {
        const cx<T> cN = coarse_at(cr, gc, (z >> 1) + 1, m.gi, m.gj[0]);
        e = (T)0.5 * (cK + cN);
        cK = cN;
      } else {
        e = cK;
      }
      xa0 = xv + e;
      ra[z & 1][qq] = xa0;
      if (z >= m.z0 && z < m.z1 && m.own[0]) yr[(int64_t)z * plane + own_off<H>(m, 0, n1)] = xa0;
    }
    const int p = z - 1;
    if (p >= max(0, m.z0 - 1) && p < m.z1 && m.in1[0]) {  // r = b - A x' on plane p
      const cx<T>* a = ra[p & 1];
      const bool lo = p == 0, hi = p == n3 - 1;
      const cx<T> xm = lo ? xa0 : xa2, xp = hi ? xa2 : xa0;
      rr[p & 1][qq] = bp - apply7(xa1, a[qq - 1], a[qq + 1], a[qq - W], a[qq + W], xm, xp, ih1, ih2, ih3, dp);
    }
    const int q = z - 2;
    if (q >= max(0, m.z0 - 1) && q < m.z1 && cown) {  // full weighting of plane q's r
      const cx<T>* rq = rr[q & 1];
      cx<T> sq = mk<T>(0, 0);
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int aa = 0; aa < 3; ++aa) sq = sq + (w[bb] * w[aa]) * rq[(lj - 1 + bb) * W + (lane - 1 + aa)];
      if (q & 1) {  // coarse plane K = (q - 1) / 2 is complete
        const int K = (q - 1) >> 1;
        if (2 * K >= m.z0) {
          const cx<T> sm = K == 0 ? sq : s2;  // plane -1 mirrors plane 1
          bc[(int64_t)K * cplane + ccol] = w[0] * sm + w[1] * s1 + w[2] * sq;
        }
      } else if (q == n3 - 1) {  // the last coarse plane: plane n3 mirrors n3 - 2
        bc[(int64_t)(q >> 1) * cplane + ccol] = w[0] * s1 + w[1] * sq + w[2] * s1;
      }
      s2 = s1;
      s1 = sq;
    }
    xa2 = xa1;
    xa1 = xa0;
    __syncthreads();
  }
}

bool fused3_ok(const Grid& gf) {
  return gf.dim == 3 && gf.bc == 2 && gf.n1 >= 16 && gf.n2 >= 16 && gf.n3 >= 3;
}

namespace {
// z-runs: at most 48 output planes each (193 x 193 x 97 at 32 RHS: three
// runs of 34, legs 1-2% faster than four of 26), and more (down to 8 planes)
// while a launch has fewer than 1024 blocks (97 x 97 x 49 at 32 RHS: two runs
// of 26, legs 5-8% faster than four of 14); runs recompute two planes on
// each side.
// A level too small to give even 256 blocks that way (one per CU) takes runs
// down to 4 planes: its launches are bound by the latency of their steps,
// not by the recomputed planes (25 x 25 x 13, 32 RHS: 21 -> 12 us per down
// leg; the same on 49 x 49 x 25, 384 blocks, was slower).
struct Runs3 {
  int ytiles, zrun;
  TileMap m;
};
Runs3 runs3(const Grid& g, int nrhs, int H) {
  Runs3 r;
  const int xtiles = (g.n1 + m3_tx(H) - 1) / m3_tx(H);
  r.ytiles = (g.n2 + m3_ty(H) - 1) / m3_ty(H);
  const int64_t per_run = (int64_t)xtiles * r.ytiles * nrhs;
  int runs = (g.n3 + 47) / 48;
  int want = (int)std::min<int64_t>((1024 + per_run - 1) / per_run, std::max(1, g.n3 / 8));
  if (per_run * want < 256) want = (int)std::min<int64_t>((256 + per_run - 1) / per_run, (g.n3 + 3) / 4);
  runs = std::max(runs, want);
  r.zrun = (g.n3 + runs - 1) / runs;
  r.zrun += r.zrun & 1;
  r.m = make_tile_map(xtiles, r.ytiles * ((g.n3 + r.zrun - 1) / r.zrun), nrhs);
  return r;
}
}  // namespace

void launch_down_fused3(const Grid& gf, const Grid& gc, int nrhs, const cx<float>* b, const cx<float>* d,
                         const cx<float>* f, double relax, cx<float>* xo, cx<float>* bco, const int* skip,
                         hipStream_t s) {
  const Runs3 r = runs3(gf, nrhs, 3);
  hipLaunchKernelGGL((k_down_fused3<float, M3_NT>), dim3(r.m.T), dim3(M3_NT), 0, s, r.m, r.ytiles, r.zrun, gf, gc,
                     b, d, f, xo, bco, skip, (float)(1.0 - relax));
}

void launch_mid_fused3(const Grid& gf, const Grid& gc, int nrhs, const cx<float>* x, const cx<float>* ec,
                       const cx<float>* b, const cx<float>* d, cx<float>* xo, cx<float>* bco, const int* skip,
                       hipStream_t s) {
  const Runs3 r = runs3(gf, nrhs, 2);
  hipLaunchKernelGGL((k_mid_fused3<float, M3_NT>), dim3(r.m.T), dim3(M3_NT), 0, s, r.m, r.ytiles, r.zrun, gf, gc, x,
                     ec, b, d, xo, bco, skip);
}

void launch_up_fused3(const Grid& gf, const Grid& gc, int nrhs, const cx<float>* x, const cx<float>* ec,
                      const cx<float>* b, const cx<float>* f, double relax, cx<float>* xo, const int* skip,
                      hipStream_t s) {
  const Runs3 r = runs3(gf, nrhs, 2);
  hipLaunchKernelGGL((k_up_fused3<float, M3_NT>), dim3(r.m.T), dim3(M3_NT), 0, s, r.m, r.ytiles, r.zrun, gf, gc, x,
                     ec, b, f, xo, skip, (float)(1.0 - relax));
}

}  // namespace hh
