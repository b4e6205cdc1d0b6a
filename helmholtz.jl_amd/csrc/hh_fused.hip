// Fused multigrid legs for 2-D levels with mirror ghosts (BC = 2) and
// 2 pre- / 2 post-sweeps, LDS-tiled with halos so each leg reads its inputs
// and writes its outputs once:
//
//   down : x1 = F b, x2 = x1 + F (b - A x1), r = b - A x2, b_c = R r
//          reads b, d; writes x2 and b_c          (unfused: 4 kernels, ~9 passes)
//   mid  : x' = x + P e_c, r = b - A x', b_c = R r
//          reads x, b, d, e_c; writes x', b_c     (unfused: 3 kernels, ~6.5 passes)
//   up   : x' = x + P e_c, two Jacobi sweeps x <- x + F (b - A x)
//          reads x, b, d, e_c; writes x''         (unfused: 3 kernels, ~8.25 passes)
//
// A block owns a TX x TY tile of fine nodes (TX, TY even, so it also owns the
// coarse nodes at even positions inside it) and stages a halo of H nodes in
// LDS.  Halo positions outside the grid hold the mirrored node's values (x~_{-k}
// = x_k): the mirror extension of a field is preserved exactly by the
// operator, the smoother, and the transfers, so ghost positions evolve like
// the nodes they mirror and every formula runs unchanged on the tile.
// F = relax (d + 2/h1^2 + 2/h2^2)^-1 is the level's precomputed damped-Jacobi
// factor (one complex multiply per update instead of a complex reciprocal).
#include "hh_kernels.h"

namespace hh {

constexpr int FTX = 64;

// mirror ghost index; positions farther out than one mirror (only loaded by
// the last tile, never used for an owned output) are clamped into the grid
__device__ __forceinline__ int mirror_idx(int g, int n) {
  g = g < 0 ? -g : (g >= n ? 2 * (n - 1) - g : g);
  return g < 0 ? 0 : (g >= n ? n - 1 : g);
}

template <typename T>
__device__ __forceinline__ cx<T> lap5(cx<T> xc, cx<T> xl, cx<T> xr, cx<T> xm, cx<T> xp, T ih1, T ih2) {
  return ih1 * (((T)2 * xc) - xl - xr) + ih2 * (((T)2 * xc) - xm - xp);
}

template <typename T, int TY, int H>
struct Tile {
  static constexpr int W = FTX + 2 * H;
  static constexpr int HT = TY + 2 * H;
  static constexpr int NPIX = W * HT;
};

// Each thread owns the tile positions q = threadIdx.x + NT k (k < K) in every
// phase, so b and d live in its registers; only the x fields (and the
// residual) are staged in LDS, where neighbours are read.  Two LDS arrays per
// block keep 6-7 blocks resident per CU to overlap one block's loads with
// another's arithmetic.

// load b and d at the owned positions of a tile whose first node is
// (fx0 - H, fy0 - H)
// (and the Jacobi factor f when given)
template <typename T, int NT, int W, int NPIX, int K, int H, bool WITH_F>
__device__ __forceinline__ void load_bd(const cx<T>* __restrict__ bb, const cx<T>* __restrict__ d,
                                        const cx<T>* __restrict__ f, int fx0, int fy0, int n1, int n2,
                                        cx<T> (&rb)[K], cx<T> (&rd)[K], cx<T> (&rf)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int q = threadIdx.x + NT * k;
    if (q < NPIX) {
      const int li = q % W, lj = q / W;
      const int64_t p = (int64_t)mirror_idx(fy0 - H + lj, n2) * n1 + mirror_idx(fx0 - H + li, n1);
      rb[k] = bb[p];
      rd[k] = d[p];
      if (WITH_F) rf[k] = f[p];
    }
  }
}

// ------------------------------------------------------------------ down leg
template <typename T, int TY, int NT>
__global__ __launch_bounds__(NT) void k_down_fused(TileMap tm, Grid gf, Grid gc, const cx<T>* __restrict__ b,
                                                    const cx<T>* __restrict__ d, const cx<T>* __restrict__ f,
                                                    cx<T>* __restrict__ xo, cx<T>* __restrict__ bco,
                                                    const int* __restrict__ skip) {
  constexpr int H = 3;
  typedef Tile<T, TY, H> Tl;
  constexpr int W = Tl::W, NPIX = Tl::NPIX, K = (NPIX + NT - 1) / NT;
  __shared__ cx<T> s1[NPIX], s2[NPIX];
  int xt, yt, rhs;
  tile_of(tm, xt, yt, rhs);
  if (skip && skip[rhs]) return;
  const int n1 = gf.n1, n2 = gf.n2;
  const int fx0 = xt * FTX, fy0 = yt * TY;
  const T ih1 = (T)gf.ih1, ih2 = (T)gf.ih2;
  cx<T> rb[K], rd[K], rf[K];
  load_bd<T, NT, W, NPIX, K, H, true>(b + (int64_t)rhs * gf.N, d, f, fx0, fy0, n1, n2, rb, rd, rf);
  // x1 = F b on the whole tile
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int q = threadIdx.x + NT * k;
    if (q < NPIX) s1[q] = rf[k] * rb[k];
  }
  __syncthreads();
  // x2 on the halo-2 region
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int q = threadIdx.x + NT * k;
    const int li = q % W, lj = q / W;
    if (q < NPIX && li >= 1 && li < W - 1 && lj >= 1 && lj < Tl::HT - 1) {
      const cx<T> xc = s1[q];
      const cx<T> Ax = lap5(xc, s1[q - 1], s1[q + 1], s1[q - W], s1[q + W], ih1, ih2) + rd[k] * xc;
      s2[q] = xc + rf[k] * (rb[k] - Ax);
    }
  }
  __syncthreads();
  // residual on the halo-1 region (over x1, which is no longer read)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int q = threadIdx.x + NT * k;
    const int li = q % W, lj = q / W;
    if (q < NPIX && li >= 2 && li < W - 2 && lj >= 2 && lj < Tl::HT - 2) {
      const cx<T> xc = s2[q];
      const cx<T> Ax = lap5(xc, s2[q - 1], s2[q + 1], s2[q - W], s2[q + W], ih1, ih2) + rd[k] * xc;
      s1[q] = rb[k] - Ax;
    }
  }
  __syncthreads();
  cx<T>* xb = xo + (int64_t)rhs * gf.N;
  for (int q = threadIdx.x; q < FTX * TY; q += NT) {
    const int i = fx0 + q % FTX, j = fy0 + q / FTX;
    if (i < n1 && j < n2) xb[(int64_t)j * n1 + i] = s2[(q / FTX + H) * W + q % FTX + H];
  }
  // full weighting of r onto the owned coarse nodes
  cx<T>* cb = bco + (int64_t)rhs * gc.N;
  const T w[3] = {(T)0.25, (T)0.5, (T)0.25};
  for (int q = threadIdx.x; q < (FTX / 2) * (TY / 2); q += NT) {
    const int I = fx0 / 2 + q % (FTX / 2), J = fy0 / 2 + q / (FTX / 2);
    if (I >= gc.n1 || J >= gc.n2) continue;
    const int lc = (2 * J - fy0 + H) * W + (2 * I - fx0 + H);
    cx<T> acc = mk<T>(0, 0);
    for (int bb2 = 0; bb2 < 3; ++bb2)
      for (int a = 0; a < 3; ++a) acc = acc + ((T)1 * w[bb2] * w[a]) * s1[lc + (bb2 - 1) * W + (a - 1)];
    cb[(int64_t)J * gc.n1 + I] = acc;
  }
}

// x + P e_c on the owned positions, staged into sx.  The coarse patch the
// tile interpolates from is first copied into LDS (sc) with coalesced loads.
// Fine nodes that influence an owned output have mirrored indices in
// [max(0, fx0 - H), fx0 + FTX - 1 + H], so their coarse neighbours lie in the
// patch starting at max(0, fx0 - H) / 2; other halo positions are clamped
// into the patch (their values are never used).
template <typename T, int TY, int H>
struct Patch {
  static constexpr int PW = (FTX + 2 * H) / 2 + 2;
  static constexpr int PH = (TY + 2 * H) / 2 + 2;
};

template <typename T, int NT, int TY, int W, int NPIX, int K, int H>
__device__ __forceinline__ void load_x_prolong(const cx<T>* __restrict__ xr, const cx<T>* __restrict__ cr,
                                               const Grid& gc, int fx0, int fy0, int n1, int n2, cx<T>* sx,
                                               cx<T>* sc) {
  constexpr int PW = Patch<T, TY, H>::PW, PH = Patch<T, TY, H>::PH;
  const int cx0 = (fx0 - H > 0 ? fx0 - H : 0) >> 1, cy0 = (fy0 - H > 0 ? fy0 - H : 0) >> 1;
  cx<T> vx[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int q = threadIdx.x + NT * k;
    if (q < NPIX) {
      const int li = q % W, lj = q / W;
      vx[k] = xr[(int64_t)mirror_idx(fy0 - H + lj, n2) * n1 + mirror_idx(fx0 - H + li, n1)];
    }
  }
  for (int q = threadIdx.x; q < PW * PH; q += NT) {
    const int I = min(cx0 + q % PW, gc.n1 - 1), J = min(cy0 + q / PW, gc.n2 - 1);
    sc[q] = cr[(int64_t)J * gc.n1 + I];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int q = threadIdx.x + NT * k;
    if (q < NPIX) {
      const int li = q % W, lj = q / W;
      const int gi = mirror_idx(fx0 - H + li, n1), gj = mirror_idx(fy0 - H + lj, n2);
      const int i0 = min(max((gi >> 1) - cx0, 0), PW - 2), j0 = min(max((gj >> 1) - cy0, 0), PH - 2);
      // nodal bilinear interpolation, same terms and order as k_prolong_add
      const T wi = (gi & 1) ? (T)0.5 : (T)1, wj = (gj & 1) ? (T)0.5 : (T)1, wk = (T)1;
      const int ni = (gi & 1) ? 2 : 1, nj = (gj & 1) ? 2 : 1;
      cx<T> acc = mk<T>(0, 0);
      for (int c2 = 0; c2 < nj; ++c2)
        for (int c1 = 0; c1 < ni; ++c1) acc = acc + (wi * wj * wk) * sc[(j0 + c2) * PW + i0 + c1];
      sx[q] = vx[k] + acc;
    }
  }
}

// ------------------------------------------------------------------ middle (W)
template <typename T, int TY, int NT>
__global__ __launch_bounds__(NT) void k_mid_fused(TileMap tm, Grid gf, Grid gc, const cx<T>* __restrict__ x,
                                                   const cx<T>* __restrict__ ec, const cx<T>* __restrict__ b,
                                                   const cx<T>* __restrict__ d, cx<T>* __restrict__ xo,
                                                   cx<T>* __restrict__ bco, const int* __restrict__ skip) {
  constexpr int H = 2;
  typedef Tile<T, TY, H> Tl;
  constexpr int W = Tl::W, NPIX = Tl::NPIX, K = (NPIX + NT - 1) / NT;
  __shared__ cx<T> sx[NPIX], sr[NPIX], sc[Patch<T, TY, H>::PW * Patch<T, TY, H>::PH];
  int xt, yt, rhs;
  tile_of(tm, xt, yt, rhs);
  if (skip && skip[rhs]) return;
  const int n1 = gf.n1, n2 = gf.n2;
  const int fx0 = xt * FTX, fy0 = yt * TY;
  const T ih1 = (T)gf.ih1, ih2 = (T)gf.ih2;
  cx<T> rb[K], rd[K], rf[K];  // rf unused (no Jacobi sweep in this leg)
  load_bd<T, NT, W, NPIX, K, H, false>(b + (int64_t)rhs * gf.N, d, nullptr, fx0, fy0, n1, n2, rb, rd, rf);
  load_x_prolong<T, NT, TY, W, NPIX, K, H>(x + (int64_t)rhs * gf.N, ec + (int64_t)rhs * gc.N, gc, fx0, fy0, n1, n2, sx,
                                       sc);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int q = threadIdx.x + NT * k;
    const int li = q % W, lj = q / W;
    if (q < NPIX && li >= 1 && li < W - 1 && lj >= 1 && lj < Tl::HT - 1) {
      const cx<T> xc = sx[q];
      const cx<T> Ax = lap5(xc, sx[q - 1], sx[q + 1], sx[q - W], sx[q + W], ih1, ih2) + rd[k] * xc;
      sr[q] = rb[k] - Ax;
    }
  }
  __syncthreads();
  cx<T>* xb = xo + (int64_t)rhs * gf.N;
  for (int q = threadIdx.x; q < FTX * TY; q += NT) {
    const int i = fx0 + q % FTX, j = fy0 + q / FTX;
    if (i < n1 && j < n2) xb[(int64_t)j * n1 + i] = sx[(q / FTX + H) * W + q % FTX + H];
  }
  cx<T>* cb = bco + (int64_t)rhs * gc.N;
  const T w[3] = {(T)0.25, (T)0.5, (T)0.25};
  for (int q = threadIdx.x; q < (FTX / 2) * (TY / 2); q += NT) {
    const int I = fx0 / 2 + q % (FTX / 2), J = fy0 / 2 + q / (FTX / 2);
    if (I >= gc.n1 || J >= gc.n2) continue;
    const int lc = (2 * J - fy0 + H) * W + (2 * I - fx0 + H);
    cx<T> acc = mk<T>(0, 0);
    for (int bb2 = 0; bb2 < 3; ++bb2)
      for (int a = 0; a < 3; ++a) acc = acc + ((T)1 * w[bb2] * w[a]) * sr[lc + (bb2 - 1) * W + (a - 1)];
    cb[(int64_t)J * gc.n1 + I] = acc;
  }
}

// ------------------------------------------------------------------ up leg
template <typename T, int TY, int NT>
__global__ __launch_bounds__(NT) void k_up_fused(TileMap tm, Grid gf, Grid gc, const cx<T>* __restrict__ x,
                                                  const cx<T>* __restrict__ ec, const cx<T>* __restrict__ b,
                                                  const cx<T>* __restrict__ d, const cx<T>* __restrict__ f,
                                                  cx<T>* __restrict__ xo, const int* __restrict__ skip) {
  constexpr int H = 2;
  typedef Tile<T, TY, H> Tl;
  constexpr int W = Tl::W, NPIX = Tl::NPIX, K = (NPIX + NT - 1) / NT;
  __shared__ cx<T> sx[NPIX], s3[NPIX], sc[Patch<T, TY, H>::PW * Patch<T, TY, H>::PH];
  int xt, yt, rhs;
  tile_of(tm, xt, yt, rhs);
  if (skip && skip[rhs]) return;
  const int n1 = gf.n1, n2 = gf.n2;
  const int fx0 = xt * FTX, fy0 = yt * TY;
  const T ih1 = (T)gf.ih1, ih2 = (T)gf.ih2;
  cx<T> rb[K], rd[K], rf[K];
  load_bd<T, NT, W, NPIX, K, H, true>(b + (int64_t)rhs * gf.N, d, f, fx0, fy0, n1, n2, rb, rd, rf);
  load_x_prolong<T, NT, TY, W, NPIX, K, H>(x + (int64_t)rhs * gf.N, ec + (int64_t)rhs * gc.N, gc, fx0, fy0, n1, n2, sx,
                                       sc);
  __syncthreads();
  auto jac = [&](const cx<T>* s, int q, int k) {
    const cx<T> xc = s[q];
    const cx<T> Ax = lap5(xc, s[q - 1], s[q + 1], s[q - W], s[q + W], ih1, ih2) + rd[k] * xc;
    return xc + rf[k] * (rb[k] - Ax);
  };
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int q = threadIdx.x + NT * k;
    const int li = q % W, lj = q / W;
    if (q < NPIX && li >= 1 && li < W - 1 && lj >= 1 && lj < Tl::HT - 1) s3[q] = jac(sx, q, k);
  }
  __syncthreads();
  cx<T>* xb = xo + (int64_t)rhs * gf.N;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int q = threadIdx.x + NT * k;
    const int li = q % W, lj = q / W;
    const int i = fx0 + li - H, j = fy0 + lj - H;
    if (q < NPIX && li >= H && li < H + FTX && lj >= H && lj < H + TY && i < n1 && j < n2)
      xb[(int64_t)j * n1 + i] = jac(s3, q, k);
  }
}

// ------------------------------------------------------------------ launchers
// tile height and threads per block: complex64 64 x 32 tiles on 512 threads,
// complex128 64 x 8 on 256 (register and LDS budget)
template <typename T>
struct FusedTY {
  static constexpr int v = sizeof(T) == 4 ? 32 : 8;
  static constexpr int nt = sizeof(T) == 4 ? 512 : 256;
};

bool fused_ok(const Grid& gf) { return gf.dim == 2 && gf.bc == 2 && gf.n1 >= 16 && gf.n2 >= 16; }

template <typename T>
TileMap tile_map(const Grid& gf, int nrhs) {
  constexpr int TY = FusedTY<T>::v;
  return make_tile_map((gf.n1 + FTX - 1) / FTX, (gf.n2 + TY - 1) / TY, nrhs);
}

template <typename T>
void launch_down_fused(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* b, const cx<T>* d, const cx<T>* f,
                       cx<T>* xo, cx<T>* bco, const int* skip, hipStream_t s) {
  constexpr int TY = FusedTY<T>::v;
  const TileMap m = tile_map<T>(gf, nrhs);
  hipLaunchKernelGGL((k_down_fused<T, TY, FusedTY<T>::nt>), dim3(m.T), dim3(FusedTY<T>::nt), 0, s, m, gf, gc, b, d, f, xo, bco,
                     skip);
}

template <typename T>
void launch_mid_fused(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* x, const cx<T>* ec, const cx<T>* b,
                      const cx<T>* d, cx<T>* xo, cx<T>* bco, const int* skip, hipStream_t s) {
  constexpr int TY = FusedTY<T>::v;
  const TileMap m = tile_map<T>(gf, nrhs);
  hipLaunchKernelGGL((k_mid_fused<T, TY, FusedTY<T>::nt>), dim3(m.T), dim3(FusedTY<T>::nt), 0, s, m, gf, gc, x, ec, b, d, xo, bco,
                     skip);
}

template <typename T>
void launch_up_fused(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* x, const cx<T>* ec, const cx<T>* b,
                     const cx<T>* d, const cx<T>* f, cx<T>* xo, const int* skip, hipStream_t s) {
  constexpr int TY = FusedTY<T>::v;
  const TileMap m = tile_map<T>(gf, nrhs);
  hipLaunchKernelGGL((k_up_fused<T, TY, FusedTY<T>::nt>), dim3(m.T), dim3(FusedTY<T>::nt), 0, s, m, gf, gc, x, ec, b, d, f, xo,
                     skip);
}

#define HH_INST(T)                                                                                              \
  template void launch_down_fused<T>(const Grid&, const Grid&, int, const cx<T>*, const cx<T>*, const cx<T>*,   \
                                     cx<T>*, cx<T>*, const int*, hipStream_t);                                  \
  template void launch_mid_fused<T>(const Grid&, const Grid&, int, const cx<T>*, const cx<T>*, const cx<T>*,    \
                                    const cx<T>*, cx<T>*, cx<T>*, const int*, hipStream_t);                     \
  template void launch_up_fused<T>(const Grid&, const Grid&, int, const cx<T>*, const cx<T>*, const cx<T>*,     \
                                   const cx<T>*, const cx<T>*, cx<T>*, const int*, hipStream_t);
HH_INST(float)
HH_INST(double)
#undef HH_INST

}  // namespace hh
