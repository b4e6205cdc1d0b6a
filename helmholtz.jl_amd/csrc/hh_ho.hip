// The spread ("high-order") Helmholtz operator of GetHelmholtzOperatorHO
// (src/GetHelmholtz.jl:54-72, src/PlainNodalLaplacian.jl:54-141), matrix-free:
//
//   H x = Lap x + M (d o x),   d = -w^2 m (1 - i gamma / Re w) + Sommerfeld
//
// 2-D (beta scalar):   Lap = A2 (x) L1 + L2 (x) A1,   M = (avb2 (x) I + I (x) avb1) / 2
// 3-D (b_l, b_m):      Lap = sum over axes q of L_q applied along q to P_q x,
//                      P_q = b_l I + (1 - b_l)/2 (av along the two other axes),
//                      M = (avb1 + avb2 + avb3) / 3 along each axis
//
// with, per axis of n nodes, the tridiagonals (row i, column i + a):
//   L   = ddxCN' ddxCN / h^2-scaled: natural-Neumann second difference
//   av  = av3term(n, 1/2), avb = av3term(n, b_m), A = (1 - b_l) av + b_l I.
// The separable forms follow from G' Gs with G, Gs kron products of ddxCN and
// av3term.  One thread per node and RHS; the arithmetic is float64.
//
// Lap and M are symmetric (sums of kron products of symmetric tridiagonals), so
// the transpose is H^T x = Lap x + d o (M x): with tr = 1 the mass term takes
// the diagonal at the row's own node instead of at each neighbour.  Adjoint
// solves use it: H^H x = b <=> H^T conj(x) = conj(b).
#include "hh_kernels.h"

namespace hh {

struct HoCoef {
  // 1-D tridiagonal entries at row i of an n-node axis, offsets -1, 0, +1
  __device__ static void lap(int i, int n, double ih2, double c[3]) {
    c[0] = i > 0 ? -ih2 : 0.0;
    c[2] = i < n - 1 ? -ih2 : 0.0;
    c[1] = (i > 0 && i < n - 1) ? 2.0 * ih2 : ih2;
  }
  // av3term(n, alpha)
  __device__ static void av(int i, int n, double alpha, double c[3]) {
    const double t = 0.5 * (1.0 - alpha);
    c[0] = i > 0 ? t : 0.0;
    c[2] = i < n - 1 ? t : 0.0;
    c[1] = (i > 0 && i < n - 1) ? alpha : 0.5 + 0.5 * alpha;
  }
};

template <typename TI>
__device__ __forceinline__ cx<double> ld(const cx<TI>* __restrict__ x, int64_t p) {
  const cx<TI> v = x[p];
  return mk<double>((double)v.re, (double)v.im);
}

// MODE: M_APPLY y = H x (+ sfac m x), M_RESID y = b - H x; tr = 1: H^T in place of H
template <typename TI, int MODE>
__global__ __launch_bounds__(256) void k_ho_stencil(Grid g, double bl, double bm, const cx<TI>* __restrict__ x,
                                                    const cx<double>* __restrict__ b, cx<double>* __restrict__ y,
                                                    const cx<double>* __restrict__ d, const double* __restrict__ m,
                                                    cx<double> sfac, const int* __restrict__ skip, int tr) {
  const int rhs = blockIdx.y;
  if (skip && skip[rhs]) return;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.N) return;
  const int64_t off = (int64_t)rhs * g.N;
  const cx<TI>* __restrict__ xr = x + off;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  const int i = (int)(p % n1);
  const int64_t q = p / n1;
  const int j = (int)(q % n2), k = (int)(q / n2);
  const int64_t s1 = 1, s2 = n1, s3 = (int64_t)n1 * n2;
  const cx<double> dp = d[p];
  cx<double> acc = mk<double>(0.0, 0.0);
  if (g.dim == 2) {
    double L1[3], L2[3], A1[3], A2[3], V1[3], V2[3];
    HoCoef::lap(i, n1, g.ih1, L1);
    HoCoef::lap(j, n2, g.ih2, L2);
    double a1[3], a2[3];
    HoCoef::av(i, n1, 0.5, a1);
    HoCoef::av(j, n2, 0.5, a2);
    for (int t = 0; t < 3; ++t) {
      A1[t] = (1.0 - bl) * a1[t] + (t == 1 ? bl : 0.0);
      A2[t] = (1.0 - bl) * a2[t] + (t == 1 ? bl : 0.0);
    }
    HoCoef::av(i, n1, bm, V1);
    HoCoef::av(j, n2, bm, V2);
    for (int bj = -1; bj <= 1; ++bj) {
      if ((bj < 0 && j == 0) || (bj > 0 && j == n2 - 1)) continue;
      for (int ai = -1; ai <= 1; ++ai) {
        if ((ai < 0 && i == 0) || (ai > 0 && i == n1 - 1)) continue;
        const int64_t pq = p + ai * s1 + bj * s2;
        const cx<double> xv = ld(xr, pq);
        const double c = A2[bj + 1] * L1[ai + 1] + L2[bj + 1] * A1[ai + 1];
        double cm = 0.0;
        if (ai == 0) cm += 0.5 * V2[bj + 1];
        if (bj == 0) cm += 0.5 * V1[ai + 1];
        acc = acc + c * xv;
        if (cm != 0.0) acc = acc + cm * ((tr ? dp : d[pq]) * xv);
      }
    }
  } else {
    const int nn[3] = {n1, n2, n3};
    const int ii[3] = {i, j, k};
    const int64_t st[3] = {s1, s2, s3};
    const double ih[3] = {g.ih1, g.ih2, g.ih3};
    double L[3][3], AV[3][3], VB[3][3];
    for (int a = 0; a < 3; ++a) {
      HoCoef::lap(ii[a], nn[a], ih[a], L[a]);
      HoCoef::av(ii[a], nn[a], 0.5, AV[a]);
      HoCoef::av(ii[a], nn[a], bm, VB[a]);
    }
    // Laplacian: for each axis q, L_q along q of P_q x
    for (int ax = 0; ax < 3; ++ax) {
      const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
      for (int u = -1; u <= 1; ++u) {
        if ((u < 0 && ii[ax] == 0) || (u > 0 && ii[ax] == nn[ax] - 1)) continue;
        const double lq = L[ax][u + 1];
        const int64_t pu = p + u * st[ax];
        // P x at pu: bl x + (1 - bl)/2 (av along o1 + av along o2)
        cx<double> px = bl * ld(xr, pu);
        for (int v = -1; v <= 1; ++v) {
          if (!((v < 0 && ii[o1] == 0) || (v > 0 && ii[o1] == nn[o1] - 1)))
            px = px + (0.5 * (1.0 - bl) * AV[o1][v + 1]) * ld(xr, pu + v * st[o1]);
          if (!((v < 0 && ii[o2] == 0) || (v > 0 && ii[o2] == nn[o2] - 1)))
            px = px + (0.5 * (1.0 - bl) * AV[o2][v + 1]) * ld(xr, pu + v * st[o2]);
        }
        acc = acc + lq * px;
      }
      // mass: (1/3) avb along this axis of d o x
      for (int u = -1; u <= 1; ++u) {
        if ((u < 0 && ii[ax] == 0) || (u > 0 && ii[ax] == nn[ax] - 1)) continue;
        const int64_t pu = p + u * st[ax];
        acc = acc + ((1.0 / 3.0) * VB[ax][u + 1]) * ((tr ? dp : d[pu]) * ld(xr, pu));
      }
    }
  }
  if (m) acc = acc + (m[p] * sfac) * ld(xr, p);
  if (MODE == M_RESID)
    y[off + p] = b[off + p] - acc;
  else
    y[off + p] = acc;
}

template <typename TI, int MODE>
void launch_ho_stencil(const Grid& g, double bl, double bm, int nrhs, const cx<TI>* x, const cx<double>* b,
                       cx<double>* y, const cx<double>* d, const double* m, cx<double> sfac, const int* skip,
                       hipStream_t s, int tr) {
  dim3 grid((unsigned)((g.N + 255) / 256), nrhs);
  hipLaunchKernelGGL((k_ho_stencil<TI, MODE>), grid, dim3(256), 0, s, g, bl, bm, x, b, y, d, m, sfac, skip, tr);
}
template void launch_ho_stencil<double, M_APPLY>(const Grid&, double, double, int, const cx<double>*,
                                                 const cx<double>*, cx<double>*, const cx<double>*, const double*,
                                                 cx<double>, const int*, hipStream_t, int);
template void launch_ho_stencil<float, M_APPLY>(const Grid&, double, double, int, const cx<float>*,
                                                const cx<double>*, cx<double>*, const cx<double>*, const double*,
                                                cx<double>, const int*, hipStream_t, int);
template void launch_ho_stencil<double, M_RESID>(const Grid&, double, double, int, const cx<double>*,
                                                 const cx<double>*, cx<double>*, const cx<double>*, const double*,
                                                 cx<double>, const int*, hipStream_t, int);

}  // namespace hh
