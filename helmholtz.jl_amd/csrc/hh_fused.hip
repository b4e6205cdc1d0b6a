// Fused multigrid legs for 2-D levels with mirror ghosts (BC = 2) and
// 2 pre- / 2 post-sweeps, LDS-tiled with halos so each leg reads its inputs
// and writes its outputs once:
//
//   down : x1 = F b, x2 = x1 + F (b - A x1), r = b - A x2, b_c = R r
//          reads b, d, F; writes b_c, and x2 only where a single coarse
//          correction follows (V levels)         (unfused: 4 kernels, ~9 passes)
//   mid  : W levels: x2 recomputed from b exactly as the down leg forms it,
//          x' = x2 + P e_c, r = b - A x', b_c = R r
//          reads b, d, F, e_c; writes x', b_c    (unfused: 3 kernels, ~6.5 passes)
//          (recomputing x2 costs ~40 flops per node and saves storing and
//          re-reading it: 16 of the 72 bytes per node and RHS of a fine-level
//          W visit)
//   up   : x' = x + P e_c, two Jacobi sweeps x <- x + F (b - A x)
//          reads x, b, d, e_c; writes x''         (unfused: 3 kernels, ~8.25 passes)
//
// A block stages a 64 x HT tile (owned FTX = 64 - 2H columns and TY = HT - 2H
// rows plus a halo of H nodes) in LDS.  One tile row is one wavefront: lane l
// always holds column l, and wavefront w rows w, w + NW, ..., so the column
// index, its mirror and the column tests are per-thread constants and the row
// index, its mirror and the row tests are wave-uniform (scalar) values.  FTX
// and TY are even, so a tile also owns the coarse nodes at even positions
// inside it.  Halo positions outside the grid hold the mirrored node's values
// (x~_{-k} = x_k): the mirror extension of a field is preserved exactly by the
// operator, the smoother, and the transfers, so ghost positions evolve like
// the nodes they mirror and every formula runs unchanged on the tile.
// F = relax D^-1, D = d + 2/h1^2 + 2/h2^2, is the level's precomputed
// damped-Jacobi factor (one complex multiply per update instead of a complex
// reciprocal).  A sweep is written with A x = D x_c - ih1 (x_l + x_r) -
// ih2 (x_m + x_p) and F D = relax as
//   x + F (b - A x) = (1 - relax) x + F (b + ih1 (x_l + x_r) + ih2 (x_m + x_p)):
// 14 floating-point operations per node instead of 26, no diagonal needed
// (the up leg reads none), the same cycle to complex64 rounding.  Residuals
// keep the operator's own form (res_at).
#include "hh_kernels.h"

namespace hh {

// tile rows: 40 (complex64, 512 threads = 8 wavefronts, 5 rows each) or
// 16 (complex128, 256 threads = 4 wavefronts, 4 rows each)
template <typename T>
struct FusedCfg {
  static constexpr int NT = sizeof(T) == 4 ? 512 : 256;
  static constexpr int HT = sizeof(T) == 4 ? 40 : 16;
};

template <typename T, int H>
struct Geo {
  static constexpr int NT = FusedCfg<T>::NT, HT = FusedCfg<T>::HT;
  static constexpr int NW = NT / 64, KR = HT / NW;
  static constexpr int FTX = 64 - 2 * H, TY = HT - 2 * H;
  static constexpr int NPIX = 64 * HT;
  // coarse patch read by the prolongation (mirrored fine indices that matter
  // lie in [max(0, fx0 - H), fx0 + FTX - 1 + H])
  static constexpr int PW = 64 / 2 + 2, PH = HT / 2 + 2;
  static_assert(HT % NW == 0, "rows per wavefront");
  static_assert(FTX % 2 == 0 && TY % 2 == 0, "owned tile must be even");
  static_assert(PW * PH <= NPIX, "coarse patch must fit in one tile array");
  static_assert(PW * PH >= NT, "every thread stores a patch value");
};

// Per-block tile coordinates and per-thread column data
struct TileCtx {
  int rhs, fx0, fy0, lane, wv;
  int gi;  // mirrored grid column of this lane
};

template <int H, int FTX, int TY>
__device__ __forceinline__ bool tile_ctx(const TileMap& tm, const Grid& gf, const int* skip, TileCtx& c) {
  int xt, yt;
  tile_of(tm, xt, yt, c.rhs);
  if (skip && skip[c.rhs]) return false;
  c.fx0 = xt * FTX;
  c.fy0 = yt * TY;
  c.lane = threadIdx.x & 63;
  c.wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  c.gi = mirror_idx(c.fx0 - H + c.lane, gf.n1);
  return true;
}

// mirrored grid row of tile row lj (wave-uniform)
template <int H>
__device__ __forceinline__ int row_of(const TileCtx& c, int lj, int n2) {
  return mirror_idx(c.fy0 - H + lj, n2);
}

// load the tile rows of this thread: v[k] = field at (row wv + NW k, column lane)
template <typename T, int H, int NW, int KR>
__device__ __forceinline__ void load_rows(const cx<T>* __restrict__ fld, const TileCtx& c, const Grid& gf,
                                          cx<T> (&v)[KR]) {
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const cx<T>* __restrict__ row = fld + (int64_t)row_of<H>(c, c.wv + NW * k, gf.n2) * gf.n1;
    v[k] = row[c.gi];
  }
}

// The coarse patch a block's prolongation reads: PW x PH values, NP per
// thread, loaded with coalesced loads into registers (load_patch) — the
// middle leg issues them with its first loads, so their latency hides behind
// its two smoothing phases — and copied into LDS (sc) where it is used.
// Positions whose values never reach an owned output are clamped into the
// patch.
template <typename T>
struct Patch {  // two values per thread (PW x PH <= 2 NT in both tile shapes)
  cx<T> v0, v1;
};

template <typename T, int H>
__device__ __forceinline__ cx<T> patch_value(const cx<T>* __restrict__ cr, const Grid& gc, int cx0, int cy0, int q) {
  constexpr int PW = Geo<T, H>::PW, PH = Geo<T, H>::PH;
  const int qq = q < PW * PH ? q : 0;
  const int I = min(cx0 + qq % PW, gc.n1 - 1), J = min(cy0 + qq / PW, gc.n2 - 1);
  return cr[(int64_t)J * gc.n1 + I];
}

template <typename T, int H>
__device__ __forceinline__ Patch<T> load_patch(const cx<T>* __restrict__ cr, const Grid& gc, const TileCtx& c) {
  typedef Geo<T, H> G;
  static_assert(G::PW * G::PH <= 2 * G::NT, "two patch values per thread");
  const int cx0 = (c.fx0 - H > 0 ? c.fx0 - H : 0) >> 1, cy0 = (c.fy0 - H > 0 ? c.fy0 - H : 0) >> 1;
  Patch<T> p;
  p.v0 = patch_value<T, H>(cr, gc, cx0, cy0, threadIdx.x);
  p.v1 = patch_value<T, H>(cr, gc, cx0, cy0, threadIdx.x + G::NT);
  return p;
}

// x + P e_c on this thread's positions, staged into sx, from the patch in
// registers (copied into sc first).  Bilinear weights per axis are (1) at
// even fine indices and (1/2, 1/2) at odd ones; the even case adds a
// zero-weighted second term so all lanes run the same code.
template <typename T, int H>
__device__ __forceinline__ void prolong_patch(const Patch<T>& pv, const Grid& gf, const TileCtx& c,
                                              const cx<T> (&vx)[Geo<T, H>::KR], cx<T>* sx, cx<T>* sc) {
  typedef Geo<T, H> G;
  constexpr int PW = G::PW, PH = G::PH;
  const int cx0 = (c.fx0 - H > 0 ? c.fx0 - H : 0) >> 1, cy0 = (c.fy0 - H > 0 ? c.fy0 - H : 0) >> 1;
  sc[threadIdx.x] = pv.v0;  // PW x PH >= NT
  if (threadIdx.x + G::NT < PW * PH) sc[threadIdx.x + G::NT] = pv.v1;
  __syncthreads();
  const int i0 = min(max((c.gi >> 1) - cx0, 0), PW - 2);
  const T wi0 = (c.gi & 1) ? (T)0.5 : (T)1, wi1 = (c.gi & 1) ? (T)0.5 : (T)0;
#pragma unroll
  for (int k = 0; k < G::KR; ++k) {
    const int lj = c.wv + G::NW * k;
    const int gj = row_of<H>(c, lj, gf.n2);
    const int j0 = min(max((gj >> 1) - cy0, 0), PH - 2);
    const cx<T>* s0 = sc + j0 * PW + i0;
    cx<T> acc = mk<T>(0, 0);
    if (gj & 1) {
      acc = acc + ((T)0.5 * wi0) * s0[0];
      acc = acc + ((T)0.5 * wi1) * s0[1];
      acc = acc + ((T)0.5 * wi0) * s0[PW];
      acc = acc + ((T)0.5 * wi1) * s0[PW + 1];
    } else {
      acc = acc + wi0 * s0[0];
      acc = acc + wi1 * s0[1];
    }
    sx[lj * 64 + c.lane] = vx[k] + acc;
  }
}

// full weighting of the residual tile sr onto the owned coarse nodes
template <typename T, int H>
__device__ __forceinline__ void restrict_tile(const cx<T>* sr, const TileCtx& c, const Grid& gc,
                                              cx<T>* __restrict__ bco) {
  typedef Geo<T, H> G;
  cx<T>* cb = bco + (int64_t)c.rhs * gc.N;
  const T w[3] = {(T)0.25, (T)0.5, (T)0.25};
  for (int q = threadIdx.x; q < (G::FTX / 2) * (G::TY / 2); q += G::NT) {
    const int I = c.fx0 / 2 + q % (G::FTX / 2), J = c.fy0 / 2 + q / (G::FTX / 2);
    if (I >= gc.n1 || J >= gc.n2) continue;
    const int lc = (2 * J - c.fy0 + H) * 64 + (2 * I - c.fx0 + H);
    cx<T> acc = mk<T>(0, 0);
    for (int bb2 = 0; bb2 < 3; ++bb2)
      for (int a = 0; a < 3; ++a) acc = acc + ((T)1 * w[bb2] * w[a]) * sr[lc + (bb2 - 1) * 64 + (a - 1)];
    cb[(int64_t)J * gc.n1 + I] = acc;
  }
}

// b + ih1 (x_l + x_r) + ih2 (x_m + x_p) at tile position q (row stride 64)
template <typename T>
__device__ __forceinline__ cx<T> rhs_nb(const cx<T>* s, int q, cx<T> bq, T ih1, T ih2) {
  const cx<T> xl = s[q - 1], xr = s[q + 1], xm = s[q - 64], xp = s[q + 64];
  return mk<T>(fma(ih1, xl.re + xr.re, fma(ih2, xm.re + xp.re, bq.re)),
               fma(ih1, xl.im + xr.im, fma(ih2, xm.im + xp.im, bq.im)));
}
// damped-Jacobi sweep at q: (1 - relax) x + F (b + ih1 sx + ih2 sy)
template <typename T>
__device__ __forceinline__ cx<T> jac_at(const cx<T>* s, int q, cx<T> bq, cx<T> fq, T ih1, T ih2, T omr) {
  const cx<T> t = rhs_nb(s, q, bq, ih1, ih2), xc = s[q];
  return mk<T>(fma(omr, xc.re, fma(fq.re, t.re, -(fq.im * t.im))), fma(omr, xc.im, fma(fq.re, t.im, fq.im * t.re)));
}
// residual at q: b - A x with the operator's own form.  The split form
// b + ih1 sx + ih2 sy - D x (or with 2 ih1 + 2 ih2 as one rounded constant)
// cancels terms of size 4 |x| / h^2 down to the small residual and perturbs
// the diagonal by an ulp of 4 / h^2 against |d| ~ k^2: it moved the complex64
// cycle 3-100x further from the exact one (up to 1e-5) where this form stays
// at 1e-7.  The sweep's split form adds no measurable error.
template <typename T>
__device__ __forceinline__ cx<T> res_at(const cx<T>* s, int q, cx<T> bq, cx<T> dq, T ih1, T ih2) {
  return bq - apply5(s[q], s[q - 1], s[q + 1], s[q - 64], s[q + 64], ih1, ih2, dq);
}
// ------------------------------------------------------------------ down leg
// STOREX: store the smoothed iterate x2 (V levels); on W levels the middle
// leg recomputes it, so only b_c is written.  A block smooths R right-hand
// sides of one tile in turn: the diagonal and the factor are loaded once and
// every RHS's b with the first loads, so the R tiles share one load latency
// (the leg waits at its barriers, not on VALU or bytes).
template <typename T, bool STOREX>
__device__ __forceinline__ void down_tile(const TileCtx& c, const Grid& gf, const Grid& gc, int rhs,
                                          const cx<T> (&rb)[Geo<T, 3>::KR], const cx<T> (&rd)[Geo<T, 3>::KR],
                                          const cx<T> (&rf)[Geo<T, 3>::KR], cx<T>* __restrict__ xo,
                                          cx<T>* __restrict__ bco, T omr, cx<T>* s1, cx<T>* s2) {
  constexpr int H = 3;
  typedef Geo<T, H> G;
  constexpr int NW = G::NW, KR = G::KR, HT = G::HT;
  const T ih1 = (T)gf.ih1, ih2 = (T)gf.ih2;
  // x1 = F b on the whole tile
#pragma unroll
  for (int k = 0; k < KR; ++k) s1[(c.wv + NW * k) * 64 + c.lane] = cmul_fma(rf[k], rb[k]);
  __syncthreads();
  // x2 on the halo-2 region; owned nodes are stored
  const bool in1 = c.lane >= 1 && c.lane < 63, own_c = c.lane >= H && c.lane < H + G::FTX && c.fx0 - H + c.lane < gf.n1;
  cx<T>* xb = xo + (int64_t)rhs * gf.N;
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int lj = c.wv + NW * k, q = lj * 64 + c.lane;
    if (lj >= 1 && lj < HT - 1 && in1) {
      const cx<T> v = jac_at(s1, q, rb[k], rf[k], ih1, ih2, omr);
      s2[q] = v;
      const int j = c.fy0 - H + lj;
      if (STOREX && lj >= H && lj < H + G::TY && j < gf.n2 && own_c) xb[(int64_t)j * gf.n1 + (c.fx0 - H + c.lane)] = v;
    }
  }
  __syncthreads();
  // residual on the halo-1 region (over x1, which is no longer read)
  const bool in2 = c.lane >= 2 && c.lane < 62;
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int lj = c.wv + NW * k, q = lj * 64 + c.lane;
    if (lj >= 2 && lj < HT - 2 && in2) s1[q] = res_at(s2, q, rb[k], rd[k], ih1, ih2);
  }
  __syncthreads();
  TileCtx cr = c;
  cr.rhs = rhs;
  restrict_tile<T, H>(s1, cr, gc, bco);
}

template <typename T, bool STOREX, int R>
__global__ __launch_bounds__(FusedCfg<T>::NT) void k_down_fused(TileMap tm, Grid gf, Grid gc,
                                                                const cx<T>* __restrict__ b,
                                                                const cx<T>* __restrict__ d,
                                                                const cx<T>* __restrict__ f, cx<T>* __restrict__ xo,
                                                                cx<T>* __restrict__ bco,
                                                                const int* __restrict__ skip, T omr, int nrhs) {
  constexpr int H = 3;
  typedef Geo<T, H> G;
  constexpr int NW = G::NW, KR = G::KR;
  __shared__ cx<T> s1[G::NPIX], s2[G::NPIX];
  TileCtx c;
  tile_ctx<H, G::FTX, G::TY>(tm, gf, nullptr, c);  // c.rhs: the block's group of R RHS
  bool act[R];
  bool any = false;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int rhs = c.rhs * R + r;
    act[r] = rhs < nrhs && !(skip && skip[rhs]);
    any = any || act[r];
  }
  if (!any) return;
  cx<T> rb[R][KR], rd[KR], rf[KR];
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (act[r]) load_rows<T, H, NW, KR>(b + (int64_t)(c.rhs * R + r) * gf.N, c, gf, rb[r]);
  load_rows<T, H, NW, KR>(d, c, gf, rd);
  load_rows<T, H, NW, KR>(f, c, gf, rf);
  bool first = true;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (!act[r]) continue;
    if (!first) __syncthreads();  // the previous RHS's full weighting has read s1
    first = false;
    down_tile<T, STOREX>(c, gf, gc, c.rhs * R + r, rb[r], rd, rf, xo, bco, omr, s1, s2);
  }
}

// ------------------------------------------------------------------ middle (W)
// x2 = F b + F (b - A F b) is formed again on the halo-2 region of an H = 3
// tile with the down leg's expressions (so it equals the x2 that leg would
// have stored, bit for bit), then x' = x2 + P e_c, the residual on the halo-1
// region, and its full weighting.  R right-hand sides per block, as in the
// down leg.
template <typename T>
__device__ __forceinline__ void mid_tile(const TileCtx& c, const Grid& gf, const Grid& gc, int rhs,
                                         const cx<T> (&rb)[Geo<T, 3>::KR], const cx<T> (&rd)[Geo<T, 3>::KR],
                                         const cx<T> (&rf)[Geo<T, 3>::KR], const Patch<T>& pv,
                                         cx<T>* __restrict__ xo, cx<T>* __restrict__ bco, T omr, cx<T>* s1,
                                         cx<T>* s2) {
  constexpr int H = 3;
  typedef Geo<T, H> G;
  constexpr int NW = G::NW, KR = G::KR, HT = G::HT;
  const T ih1 = (T)gf.ih1, ih2 = (T)gf.ih2;
  cx<T> vx[KR];
  // x1 = F b on the whole tile
#pragma unroll
  for (int k = 0; k < KR; ++k) s1[(c.wv + NW * k) * 64 + c.lane] = cmul_fma(rf[k], rb[k]);
  __syncthreads();
  // x2 on the halo-2 region (the down leg's second sweep)
  const bool in1 = c.lane >= 1 && c.lane < 63;
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int lj = c.wv + NW * k, q = lj * 64 + c.lane;
    if (lj >= 1 && lj < HT - 1 && in1) s2[q] = jac_at(s1, q, rb[k], rf[k], ih1, ih2, omr);
  }
  __syncthreads();
  // x2 into registers (the outer ring holds no x2; its x' never reaches an
  // owned output), then the coarse patch may overwrite s2
#pragma unroll
  for (int k = 0; k < KR; ++k) vx[k] = s2[(c.wv + NW * k) * 64 + c.lane];
  __syncthreads();
  prolong_patch<T, H>(pv, gf, c, vx, s1, s2);
  __syncthreads();
  // residual on the halo-1 region over x' (s1); owned x' is stored
  const bool in2 = c.lane >= 2 && c.lane < 62, own_c = c.lane >= H && c.lane < H + G::FTX && c.fx0 - H + c.lane < gf.n1;
  cx<T>* xb = xo + (int64_t)rhs * gf.N;
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int lj = c.wv + NW * k, q = lj * 64 + c.lane;
    if (lj >= 2 && lj < HT - 2 && in2) s2[q] = res_at(s1, q, rb[k], rd[k], ih1, ih2);
    const int j = c.fy0 - H + lj;
    if (lj >= H && lj < H + G::TY && j < gf.n2 && own_c) xb[(int64_t)j * gf.n1 + (c.fx0 - H + c.lane)] = s1[q];
  }
  __syncthreads();
  TileCtx cr = c;
  cr.rhs = rhs;
  restrict_tile<T, H>(s2, cr, gc, bco);
}

template <typename T, int R>
__global__ __launch_bounds__(FusedCfg<T>::NT) void k_mid_fused(TileMap tm, Grid gf, Grid gc,
                                                               const cx<T>* __restrict__ ec,
                                                               const cx<T>* __restrict__ b,
                                                               const cx<T>* __restrict__ d,
                                                               const cx<T>* __restrict__ f, cx<T>* __restrict__ xo,
                                                               cx<T>* __restrict__ bco,
                                                               const int* __restrict__ skip, T omr, int nrhs) {
  constexpr int H = 3;
  typedef Geo<T, H> G;
  constexpr int NW = G::NW, KR = G::KR;
  __shared__ cx<T> s1[G::NPIX], s2[G::NPIX];
  TileCtx c;
  tile_ctx<H, G::FTX, G::TY>(tm, gf, nullptr, c);  // c.rhs: the block's group of R RHS
  bool act[R];
  bool any = false;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int rhs = c.rhs * R + r;
    act[r] = rhs < nrhs && !(skip && skip[rhs]);
    any = any || act[r];
  }
  if (!any) return;
  cx<T> rb[R][KR], rd[KR], rf[KR];
  Patch<T> pv[R];
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (act[r]) {
      load_rows<T, H, NW, KR>(b + (int64_t)(c.rhs * R + r) * gf.N, c, gf, rb[r]);
      pv[r] = load_patch<T, H>(ec + (int64_t)(c.rhs * R + r) * gc.N, gc, c);
    }
  load_rows<T, H, NW, KR>(d, c, gf, rd);
  load_rows<T, H, NW, KR>(f, c, gf, rf);
  bool first = true;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (!act[r]) continue;
    if (!first) __syncthreads();  // the previous RHS's full weighting has read s2
    first = false;
    mid_tile<T>(c, gf, gc, c.rhs * R + r, rb[r], rd, rf, pv[r], xo, bco, omr, s1, s2);
  }
}

// ------------------------------------------------------------------ up leg
template <typename T>
__global__ __launch_bounds__(FusedCfg<T>::NT) void k_up_fused(TileMap tm, Grid gf, Grid gc,
                                                              const cx<T>* __restrict__ x,
                                                              const cx<T>* __restrict__ ec,
                                                              const cx<T>* __restrict__ b,
                                                              const cx<T>* __restrict__ f, cx<T>* __restrict__ xo,
                                                              const int* __restrict__ skip, T omr) {
  constexpr int H = 2;
  typedef Geo<T, H> G;
  constexpr int NW = G::NW, KR = G::KR, HT = G::HT;
  // the coarse patch is staged in s3, which is first written after the
  // prolongation's last read of the patch (barrier between)
  __shared__ cx<T> sx[G::NPIX], s3[G::NPIX];
  TileCtx c;
  if (!tile_ctx<H, G::FTX, G::TY>(tm, gf, skip, c)) return;
  const T ih1 = (T)gf.ih1, ih2 = (T)gf.ih2;
  const int64_t off = (int64_t)c.rhs * gf.N;
  cx<T> rb[KR], rf[KR], vx[KR];
  load_rows<T, H, NW, KR>(x + off, c, gf, vx);
  load_rows<T, H, NW, KR>(b + off, c, gf, rb);
  load_rows<T, H, NW, KR>(f, c, gf, rf);
  const Patch<T> pv = load_patch<T, H>(ec + (int64_t)c.rhs * gc.N, gc, c);
  prolong_patch<T, H>(pv, gf, c, vx, sx, s3);
  __syncthreads();
  const bool in1 = c.lane >= 1 && c.lane < 63, own_c = c.lane >= H && c.lane < H + G::FTX && c.fx0 - H + c.lane < gf.n1;
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int lj = c.wv + NW * k, q = lj * 64 + c.lane;
    if (lj >= 1 && lj < HT - 1 && in1) s3[q] = jac_at(sx, q, rb[k], rf[k], ih1, ih2, omr);
  }
  __syncthreads();
  cx<T>* xb = xo + off;
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int lj = c.wv + NW * k, q = lj * 64 + c.lane;
    const int j = c.fy0 - H + lj;
    if (lj >= H && lj < H + G::TY && j < gf.n2 && own_c)
      xb[(int64_t)j * gf.n1 + (c.fx0 - H + c.lane)] = jac_at(s3, q, rb[k], rf[k], ih1, ih2, omr);
  }
}

// ------------------------------------------------------------------ launchers
bool fused_ok(const Grid& gf) { return gf.dim == 2 && gf.bc == 2 && gf.n1 >= 16 && gf.n2 >= 16; }

template <typename T, int H>
TileMap tile_map(const Grid& gf, int nrhs) {
  typedef Geo<T, H> G;
  return make_tile_map((gf.n1 + G::FTX - 1) / G::FTX, (gf.n2 + G::TY - 1) / G::TY, nrhs);
}

// Right-hand sides per block: 2 in the down leg (62 VGPRs, eight wavefronts
// per SIMD still fit), 1 in the middle leg (two RHS need 104 VGPRs and halve
// the resident blocks: 0.81 against 0.73 ms on the 1025 x 513 grid).
template <typename T, int R>
static void down_r_launch(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* b, const cx<T>* d, const cx<T>* f,
                          T omr, cx<T>* xo, cx<T>* bco, const int* skip, hipStream_t s) {
  const TileMap m = tile_map<T, 3>(gf, (nrhs + R - 1) / R);
  if (xo)
    hipLaunchKernelGGL((k_down_fused<T, true, R>), dim3(m.T), dim3(FusedCfg<T>::NT), 0, s, m, gf, gc, b, d, f, xo,
                       bco, skip, omr, nrhs);
  else
    hipLaunchKernelGGL((k_down_fused<T, false, R>), dim3(m.T), dim3(FusedCfg<T>::NT), 0, s, m, gf, gc, b, d, f, xo,
                       bco, skip, omr, nrhs);
}

template <typename T>
void launch_down_fused(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* b, const cx<T>* d, const cx<T>* f,
                       double relax, cx<T>* xo, cx<T>* bco, const int* skip, hipStream_t s) {
  const T omr = (T)(1.0 - relax);
  if (nrhs > 1)
    down_r_launch<T, 2>(gf, gc, nrhs, b, d, f, omr, xo, bco, skip, s);
  else
    down_r_launch<T, 1>(gf, gc, nrhs, b, d, f, omr, xo, bco, skip, s);
}

template <typename T>
void launch_mid_fused(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* ec, const cx<T>* b, const cx<T>* d,
                      const cx<T>* f, double relax, cx<T>* xo, cx<T>* bco, const int* skip, hipStream_t s) {
  const TileMap m = tile_map<T, 3>(gf, nrhs);
  hipLaunchKernelGGL((k_mid_fused<T, 1>), dim3(m.T), dim3(FusedCfg<T>::NT), 0, s, m, gf, gc, ec, b, d, f, xo, bco,
                     skip, (T)(1.0 - relax), nrhs);
}

template <typename T>
void launch_up_fused(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* x, const cx<T>* ec, const cx<T>* b,
                     const cx<T>* f, double relax, cx<T>* xo, const int* skip, hipStream_t s) {
  const TileMap m = tile_map<T, 2>(gf, nrhs);
  hipLaunchKernelGGL((k_up_fused<T>), dim3(m.T), dim3(FusedCfg<T>::NT), 0, s, m, gf, gc, x, ec, b, f, xo, skip,
                     (T)(1.0 - relax));
}

#define HH_INST(T)                                                                                              \
  template void launch_down_fused<T>(const Grid&, const Grid&, int, const cx<T>*, const cx<T>*, const cx<T>*,   \
                                     double, cx<T>*, cx<T>*, const int*, hipStream_t);                          \
  template void launch_mid_fused<T>(const Grid&, const Grid&, int, const cx<T>*, const cx<T>*, const cx<T>*,    \
                                    const cx<T>*, double, cx<T>*, cx<T>*, const int*, hipStream_t);             \
  template void launch_up_fused<T>(const Grid&, const Grid&, int, const cx<T>*, const cx<T>*, const cx<T>*,     \
                                   const cx<T>*, double, cx<T>*, const int*, hipStream_t);
HH_INST(float)
HH_INST(double)
#undef HH_INST

}  // namespace hh
