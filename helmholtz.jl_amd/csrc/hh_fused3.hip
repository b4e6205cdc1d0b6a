// Fused up leg for 3-D levels with mirror ghosts (BC = 2), complex64:
//
//   x' = x + P e_c, then two damped-Jacobi sweeps x <- x + F (b - A x)
//
// reading x, b, d, F and the coarse correction once and writing the result
// once (unfused: the prolongation and two sweeps, 3 passes).  A block owns a
// 60 x 12 column of nodes (a 64 x 16 in-plane tile with a 2-node halo, one tile
// row per wavefront as in the 2-D legs) and marches it along z.  Step z forms
// x' on plane z, the first sweep on plane z - 1 and the second on plane z - 2;
// x' and the first sweep live in LDS rings of three planes, and b, d, F of the
// three planes in flight in registers; one barrier per step.  Each block covers a run of 16 output
// planes.  Ghost planes mirror (plane -1 = 1,
// n3 = n3 - 2) like the in-plane halos, so every formula is the unfused one:
// the result equals prolongation + two sweeps.
#include "hh_kernels.h"

namespace hh {

namespace {
constexpr int U3_H = 2, U3_RY = 16, U3_NW = 8, U3_KR = U3_RY / U3_NW;
constexpr int U3_FTX = 64 - 2 * U3_H, U3_TY = U3_RY - 2 * U3_H, U3_NP = 64 * U3_RY;

}  // namespace

// Blocks also split z into runs of zrun output planes (zrun even); a run
// starting at z0 > 0 first forms x' on planes z0 - 2, z0 - 1 and the first
// sweep on z0 - 1 (recomputed, not exchanged), so runs are independent.
template <typename T>
__global__ __launch_bounds__(512) void k_up_fused3(TileMap tm, int ytiles, int zrun, Grid gf, Grid gc,
                                                   const cx<T>* __restrict__ x, const cx<T>* __restrict__ ec,
                                                   const cx<T>* __restrict__ b, const cx<T>* __restrict__ d,
                                                   const cx<T>* __restrict__ f, cx<T>* __restrict__ xo,
                                                   const int* __restrict__ skip) {
  constexpr int H = U3_H, NW = U3_NW, KR = U3_KR, RY = U3_RY;
  __shared__ cx<T> ra[3][U3_NP], rs[3][U3_NP];
  int xt, yz, rhs;
  tile_of(tm, xt, yz, rhs);
  if (skip && skip[rhs]) return;
  const int yt = yz % ytiles, zc = yz / ytiles;
  const int n1 = gf.n1, n2 = gf.n2, n3 = gf.n3;
  const int z0 = zc * zrun, z1 = min(n3, z0 + zrun);
  const int za = max(0, z0 - 2), zb = min(n3, z1 + 2);  // planes of x'
  const int fx0 = xt * U3_FTX, fy0 = yt * U3_TY;
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gi = mirror_idx(fx0 - H + lane, n1);
  const int64_t plane = (int64_t)n1 * n2, off = (int64_t)rhs * gf.N;
  const int64_t cplane = (int64_t)gc.n1 * gc.n2;
  const cx<T>* __restrict__ xr = x + off;
  const cx<T>* __restrict__ br = b + off;
  const cx<T>* __restrict__ cr = ec + (int64_t)rhs * gc.N;
  cx<T>* __restrict__ yr = xo + off;
  const T ih1 = (T)gf.ih1, ih2 = (T)gf.ih2, ih3 = (T)gf.ih3;
  // in-plane position data (rows are wave-uniform)
  int lj[KR], gj[KR];
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    lj[k] = wv + NW * k;
    gj[k] = mirror_idx(fy0 - H + lj[k], n2);
  }
  const bool in1 = lane >= 1 && lane < 63;
  const bool own_c = lane >= H && lane < H + U3_FTX && fx0 - H + lane < n1;
  // in-plane bilinear interpolant of coarse plane K at this thread's positions
  const int ci0 = gi >> 1, ci1 = (gi & 1) ? ci0 + 1 : ci0;
  const T wi = (gi & 1) ? (T)0.5 : (T)1;
  auto cval = [&](int K, int k) {
    const cx<T>* cp = cr + (int64_t)K * cplane;
    const int cj0 = gj[k] >> 1, cj1 = (gj[k] & 1) ? cj0 + 1 : cj0;
    const T w = wi * ((gj[k] & 1) ? (T)0.5 : (T)1);
    const int64_t r0 = (int64_t)cj0 * gc.n1, r1 = (int64_t)cj1 * gc.n1;
    cx<T> acc = w * cp[r0 + ci0];
    if (gi & 1) acc = acc + w * cp[r0 + ci1];
    if (gj[k] & 1) {
      acc = acc + w * cp[r1 + ci0];
      if (gi & 1) acc = acc + w * cp[r1 + ci1];
    }
    return acc;
  };
  cx<T> cK[KR];
#pragma unroll
  for (int k = 0; k < KR; ++k) cK[k] = cval(za >> 1, k);
  // b, d, F of planes z (0), z - 1 (1), z - 2 (2); this thread's own x' of
  // planes z, z - 1, z - 2 and first-sweep values of planes z - 1, z - 2,
  // z - 3 (the z-neighbours of its positions come from registers, only the
  // in-plane neighbours from LDS)
  cx<T> b0[KR], d0[KR], f0[KR], b1[KR], d1[KR], f1[KR], b2[KR], d2[KR], f2[KR];
  cx<T> xa0[KR], xa1[KR], xa2[KR], sa0[KR], sa1[KR], sa2[KR];
  // x, b, d, F of the next plane, loaded a step before they are used
  cx<T> xn[KR], bn[KR], dn[KR], fn[KR];
  auto load_plane = [&](int zz) {
    const int64_t pz = (int64_t)zz * plane;
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const int64_t o = pz + (int64_t)gj[k] * n1 + gi;
      xn[k] = xr[o];
      bn[k] = br[o];
      dn[k] = d[o];
      fn[k] = f[o];
    }
  };
  load_plane(za);
  for (int z = za; z <= z1 + 1; ++z) {
    if (z < zb) {  // x' on plane z (its x, b, d, F were loaded one step ahead)
      cx<T> xv[KR];
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        xv[k] = xn[k];
        b0[k] = bn[k];
        d0[k] = dn[k];
        f0[k] = fn[k];
      }
      if (z + 1 < zb) load_plane(z + 1);
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
        ra[z % 3][lj[k] * 64 + lane] = xa0[k];
      }
    }
    if (z - 1 >= max(0, z0 - 1) && z - 1 < min(n3, z1 + 1)) {  // first sweep on plane p = z - 1
      const int p = z - 1;
      const cx<T>* a0 = ra[p % 3];
      cx<T>* s = rs[p % 3];
      const bool lo = p == 0, hi = p == n3 - 1;  // mirrored z-neighbour
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        const int q = lj[k] * 64 + lane;
        if (lj[k] >= 1 && lj[k] < RY - 1 && in1) {
          const cx<T> xc = xa1[k];
          const cx<T> xm = lo ? xa0[k] : xa2[k], xp = hi ? xa2[k] : xa0[k];
          const cx<T> Ax = apply7(xc, a0[q - 1], a0[q + 1], a0[q - 64], a0[q + 64], xm, xp, ih1, ih2, ih3, d1[k]);
          sa0[k] = xc + f1[k] * (b1[k] - Ax);
          s[q] = sa0[k];
        }
      }
    }
    if (z - 2 >= z0 && z - 2 < z1) {  // second sweep on plane r = z - 2, owned nodes stored
      const int r = z - 2;
      const cx<T>* s0 = rs[r % 3];
      const int64_t pr = (int64_t)r * plane;
      const bool lo = r == 0, hi = r == n3 - 1;
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        const int q = lj[k] * 64 + lane;
        const int j = fy0 - H + lj[k];
        if (lj[k] >= H && lj[k] < H + U3_TY && j < n2 && own_c) {
          const cx<T> xc = sa1[k];
          const cx<T> sm = lo ? sa0[k] : sa2[k], sp = hi ? sa2[k] : sa0[k];
          const cx<T> Ax = apply7(xc, s0[q - 1], s0[q + 1], s0[q - 64], s0[q + 64], sm, sp, ih1, ih2, ih3, d2[k]);
          yr[pr + (int64_t)j * n1 + (fx0 - H + lane)] = xc + f2[k] * (b2[k] - Ax);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      b2[k] = b1[k];
      d2[k] = d1[k];
      f2[k] = f1[k];
      b1[k] = b0[k];
      d1[k] = d0[k];
      f1[k] = f0[k];
      xa2[k] = xa1[k];
      xa1[k] = xa0[k];
      sa2[k] = sa1[k];
      sa1[k] = sa0[k];
    }
    // each stage reads in LDS only planes written in earlier steps (the
    // z-neighbours it needs from this step are the thread's own registers),
    // so one barrier per step separates every write from its readers
    __syncthreads();
  }
}

bool fused3_ok(const Grid& gf) {
  return gf.dim == 3 && gf.bc == 2 && gf.n1 >= 16 && gf.n2 >= 16 && gf.n3 >= 3;
}

void launch_up_fused3(const Grid& gf, const Grid& gc, int nrhs, const cx<float>* x, const cx<float>* ec,
                      const cx<float>* b, const cx<float>* d, const cx<float>* f, cx<float>* xo, const int* skip,
                      hipStream_t s) {
  const int ytiles = (gf.n2 + U3_TY - 1) / U3_TY, zrun = 16;
  const TileMap m = make_tile_map((gf.n1 + U3_FTX - 1) / U3_FTX, ytiles * ((gf.n3 + zrun - 1) / zrun), nrhs);
  hipLaunchKernelGGL((k_up_fused3<float>), dim3(m.T), dim3(512), 0, s, m, ytiles, zrun, gf, gc, x, ec, b, d, f, xo,
                     skip);
}

}  // namespace hh
