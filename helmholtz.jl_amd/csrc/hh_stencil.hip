// Matrix-free Helmholtz stencil, damped Jacobi, grid transfers, casts and the
// dense coarsest-level solve.  All HBM-bound streaming kernels.
//
// Operator (src/GetHelmholtz.jl:33-50, src/PlainNodalLaplacian.jl:18-46):
//   (A x)_p = sum_d (2 x_p - x~_{p-e_d} - x~_{p+e_d}) / h_d^2 + d_p x_p
// with mirror ghosts (BC=2) or ghost = self (BC=1) on every face, and the
// per-node complex diagonal d_p = -w^2 m_p (1 - i g_p / Re w) + Sommerfeld
// terms (+ i s w^2 m_p for a shifted operator).
//
// Launch shape: 256 threads over consecutive positions of a row (2-D) or of
// a plane (3-D), numbered across the grid so no lane idles at a ragged row
// end; each thread marches a run of rows/planes along the slowest axis keeping
// the previous/current/next values in registers, so x is read from HBM once;
// in-plane neighbours come from the same or nearby cache lines.  The XCD-aware
// block order groups right-hand sides that share the diagonal in L2.
#include "hh_kernels.h"

namespace hh {

template <typename TD>
__device__ __forceinline__ cx<TD> node_diag(const cx<TD>* __restrict__ d, const double* __restrict__ m,
                                            cx<double> sfac, int64_t p) {
  cx<TD> dv = d[p];
  if (m) {
    double mv = m[p];
    dv.re += (TD)(sfac.re * mv);
    dv.im += (TD)(sfac.im * mv);
  }
  return dv;
}

// M_JACOBI: x + F (b - A x) with the level's precomputed damped-Jacobi factor
// F = wj (d + diag of the Laplacian)^-1 when fj is given, else wj (b - A x) / Dtot
template <typename TI, typename TO, typename TD, int MODE>
__device__ __forceinline__ void finish(cx<TO>* __restrict__ yo, const cx<TO>* __restrict__ bo, int64_t p,
                                       cx<TD> xc, cx<TD> Ax, cx<TD> Dtot, TD wj, const cx<TD>* __restrict__ fj) {
  if (MODE == M_APPLY) {
    yo[p] = cvt<TD, TO>(Ax);
  } else if (MODE == M_RESID) {
    cx<TD> bv = cvt<TO, TD>(bo[p]);
    yo[p] = cvt<TD, TO>(bv - Ax);
  } else {  // M_JACOBI
    cx<TD> bv = cvt<TO, TD>(bo[p]);
    cx<TD> r = bv - Ax;
    if (fj) {
      yo[p] = cvt<TD, TO>(xc + fj[p] * r);
    } else {
      cx<TD> up = r * cinv(Dtot);
      yo[p] = cvt<TD, TO>(xc + wj * up);
    }
  }
}

// One thread per (column position, run of `rows` rows/planes): the positions
// of a row (2-D) or of a plane (3-D) are numbered consecutively across the
// whole grid, so a block of 256 threads covers 256 consecutive nodes with no
// idle lanes at ragged row ends; each thread then marches its run along the
// slowest axis keeping the previous/current/next values in registers.
template <typename TI, typename TO, typename TD, int MODE>
__global__ __launch_bounds__(256) void k_stencil2d(Grid g, const cx<TI>* __restrict__ x,
                                                   const cx<TO>* __restrict__ b, cx<TO>* __restrict__ y,
                                                   const cx<TD>* __restrict__ d, const double* __restrict__ m,
                                                   cx<double> sfac, TD wj, int rows, TileMap tm,
                                                   const int* __restrict__ skip, const cx<TD>* __restrict__ fj) {
  int bx, by, rhs;
  tile_of(tm, bx, by, rhs);
  if (skip && skip[rhs]) return;
  const int n1 = g.n1, n2 = g.n2;
  const unsigned item = (unsigned)bx * 256u + threadIdx.x, band = item / (unsigned)n1;
  const int i = (int)(item - band * (unsigned)n1);
  const int j0 = (int)band * rows;
  if (j0 >= n2) return;
  const int j1 = min(j0 + rows, n2);
  const int64_t off = (int64_t)rhs * g.N;
  const cx<TI>* xr = x + off;
  const cx<TO>* bo = b ? b + off : nullptr;
  cx<TO>* yo = y + off;
  const int il = lo_nb(i, g.bc), ir = hi_nb(i, n1, g.bc);
  const TD ih1 = (TD)g.ih1, ih2 = (TD)g.ih2;
  const TD ldx = (TD)(g.ih1 * lap_diag1(i, n1, g.bc));
  cx<TD> xm = cvt<TI, TD>(xr[(int64_t)lo_nb(j0, g.bc) * n1 + i]);
  cx<TD> xc = cvt<TI, TD>(xr[(int64_t)j0 * n1 + i]);
  for (int j = j0; j < j1; ++j) {
    const int64_t row = (int64_t)j * n1;
    const int jp = hi_nb(j, n2, g.bc);
    cx<TD> xp = cvt<TI, TD>(xr[(int64_t)jp * n1 + i]);
    cx<TD> xl = cvt<TI, TD>(xr[row + il]);
    cx<TD> xrr = cvt<TI, TD>(xr[row + ir]);
    const int64_t p = row + i;
    cx<TD> dv = node_diag<TD>(d, m, sfac, p);
    cx<TD> Ax = apply5(xc, xl, xrr, xm, xp, ih1, ih2, dv);
    cx<TD> Dtot = dv;
    if (MODE == M_JACOBI && !fj) {
      Dtot.re += ldx + (TD)(g.ih2 * lap_diag1(j, n2, g.bc));
    }
    finish<TI, TO, TD, MODE>(yo, bo, p, xc, Ax, Dtot, wj, fj);
    xm = xc;
    xc = xp;
  }
}

template <typename TI, typename TO, typename TD, int MODE>
__global__ __launch_bounds__(256) void k_stencil3d(Grid g, const cx<TI>* __restrict__ x,
                                                   const cx<TO>* __restrict__ b, cx<TO>* __restrict__ y,
                                                   const cx<TD>* __restrict__ d, const double* __restrict__ m,
                                                   cx<double> sfac, TD wj, int rows, TileMap tm,
                                                   const int* __restrict__ skip, const cx<TD>* __restrict__ fj) {
  int bx, by, rhs;
  tile_of(tm, bx, by, rhs);
  if (skip && skip[rhs]) return;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  const unsigned item = (unsigned)bx * 256u + threadIdx.x, jj = item / (unsigned)n1;
  const int i = (int)(item - jj * (unsigned)n1), j = (int)jj;
  const int k0 = by * rows;
  if (j >= n2 || k0 >= n3) return;
  const int k1 = min(k0 + rows, n3);
  const int64_t plane = (int64_t)n1 * n2;
  const int64_t off = (int64_t)rhs * g.N;
  const cx<TI>* xr = x + off;
  const cx<TO>* bo = b ? b + off : nullptr;
  cx<TO>* yo = y + off;
  const int il = lo_nb(i, g.bc), ir = hi_nb(i, n1, g.bc);
  const int jl = lo_nb(j, g.bc), jr = hi_nb(j, n2, g.bc);
  const TD ih1 = (TD)g.ih1, ih2 = (TD)g.ih2, ih3 = (TD)g.ih3;
  const TD ldxy = (TD)(g.ih1 * lap_diag1(i, n1, g.bc) + g.ih2 * lap_diag1(j, n2, g.bc));
  const int64_t col = (int64_t)j * n1 + i;
  const int64_t c1l = (int64_t)j * n1 + il, c1r = (int64_t)j * n1 + ir;
  const int64_t c2l = (int64_t)jl * n1 + i, c2r = (int64_t)jr * n1 + i;
  cx<TD> xm = cvt<TI, TD>(xr[(int64_t)lo_nb(k0, g.bc) * plane + col]);
  cx<TD> xc = cvt<TI, TD>(xr[(int64_t)k0 * plane + col]);
  for (int k = k0; k < k1; ++k) {
    const int64_t pl = (int64_t)k * plane;
    const int kp = hi_nb(k, n3, g.bc);
    cx<TD> xp = cvt<TI, TD>(xr[(int64_t)kp * plane + col]);
    cx<TD> x1l = cvt<TI, TD>(xr[pl + c1l]);
    cx<TD> x1r = cvt<TI, TD>(xr[pl + c1r]);
    cx<TD> x2l = cvt<TI, TD>(xr[pl + c2l]);
    cx<TD> x2r = cvt<TI, TD>(xr[pl + c2r]);
    const int64_t p = pl + col;
    cx<TD> dv = node_diag<TD>(d, m, sfac, p);
    cx<TD> Ax = apply7(xc, x1l, x1r, x2l, x2r, xm, xp, ih1, ih2, ih3, dv);
    cx<TD> Dtot = dv;
    if (MODE == M_JACOBI && !fj) {
      Dtot.re += ldxy + (TD)(g.ih3 * lap_diag1(k, n3, g.bc));
    }
    finish<TI, TO, TD, MODE>(yo, bo, p, xc, Ax, Dtot, wj, fj);
    xm = xc;
    xc = xp;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_jacobi0(Grid g, const cx<T>* __restrict__ b, cx<T>* __restrict__ x,
                                                 const cx<T>* __restrict__ f, const int* __restrict__ skip) {
  const int rhs = blockIdx.y;
  if (skip && skip[rhs]) return;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.N) return;
  const int64_t o = (int64_t)rhs * g.N + p;
  x[o] = cmul_fma(f[p], b[o]);
}

// Full weighting with mirror ghosts, fine residual -> coarse right-hand side.
template <typename T>
__global__ __launch_bounds__(256) void k_restrict(Grid gf, Grid gc, const cx<T>* __restrict__ rf,
                                                  cx<T>* __restrict__ bc, const int* __restrict__ skip) {
  const int rhs = blockIdx.y;
  if (skip && skip[rhs]) return;
  const int64_t pc = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (pc >= gc.N) return;
  int I, J, K;
  node_coords(gc, pc, I, J, K);
  const cx<T>* r = rf + (int64_t)rhs * gf.N;
  const T w[3] = {(T)0.25, (T)0.5, (T)0.25};
  cx<T> acc = mk<T>(0, 0);
  const int kk = gf.dim == 3 ? 3 : 1;
  for (int c = 0; c < kk; ++c) {
    int fk = 0;
    T wk = (T)1;
    if (gf.dim == 3) {
      int f = 2 * K + c - 1;
      fk = f < 0 ? 1 : (f >= gf.n3 ? gf.n3 - 2 : f);
      wk = w[c];
    }
    for (int bb = 0; bb < 3; ++bb) {
      int f2 = 2 * J + bb - 1;
      int fj = f2 < 0 ? 1 : (f2 >= gf.n2 ? gf.n2 - 2 : f2);
      for (int a = 0; a < 3; ++a) {
        int f1 = 2 * I + a - 1;
        int fi = f1 < 0 ? 1 : (f1 >= gf.n1 ? gf.n1 - 2 : f1);
        acc = acc + (wk * w[bb] * w[a]) * r[((int64_t)fk * gf.n2 + fj) * gf.n1 + fi];
      }
    }
  }
  bc[(int64_t)rhs * gc.N + pc] = acc;
}

// 3-D full weighting marching along z: a thread owns one coarse (I, J) column
// over a run of coarse planes; the in-plane 9-point weighted sum s_f of each
// fine plane f at (I, J) is formed once and the coarse value is
// s_{2K-1} / 4 + s_{2K} / 2 + s_{2K+1} / 4 (mirror ghosts on every axis).
template <typename T>
__global__ __launch_bounds__(256) void k_restrict3(Grid gf, Grid gc, int rows, TileMap tm,
                                                   const cx<T>* __restrict__ rf, cx<T>* __restrict__ bc,
                                                   const int* __restrict__ skip) {
  int bx, by, rhs;
  tile_of(tm, bx, by, rhs);
  if (skip && skip[rhs]) return;
  const unsigned item = (unsigned)bx * 256u + threadIdx.x, JJ = item / (unsigned)gc.n1;
  const int I = (int)(item - JJ * (unsigned)gc.n1), J = (int)JJ;
  const int K0 = by * rows;
  if (J >= gc.n2 || K0 >= gc.n3) return;
  const int K1 = min(K0 + rows, gc.n3);
  const cx<T>* r = rf + (int64_t)rhs * gf.N;
  cx<T>* out = bc + (int64_t)rhs * gc.N;
  const T w[3] = {(T)0.25, (T)0.5, (T)0.25};
  int64_t cols[3][3];
  for (int bb = 0; bb < 3; ++bb) {
    const int f2 = 2 * J + bb - 1;
    const int fj = f2 < 0 ? 1 : (f2 >= gf.n2 ? gf.n2 - 2 : f2);
    for (int aa = 0; aa < 3; ++aa) {
      const int f1 = 2 * I + aa - 1;
      const int fi = f1 < 0 ? 1 : (f1 >= gf.n1 ? gf.n1 - 2 : f1);
      cols[bb][aa] = (int64_t)fj * gf.n1 + fi;
    }
  }
  const int64_t plane = (int64_t)gf.n1 * gf.n2;
  auto plane_sum = [&](int f) {
    const int fk = f < 0 ? 1 : (f >= gf.n3 ? gf.n3 - 2 : f);
    const cx<T>* rp = r + (int64_t)fk * plane;
    cx<T> acc = mk<T>(0, 0);
    for (int bb = 0; bb < 3; ++bb)
      for (int aa = 0; aa < 3; ++aa) acc = acc + (w[bb] * w[aa]) * rp[cols[bb][aa]];
    return acc;
  };
  const int64_t ccol = (int64_t)J * gc.n1 + I, cplane = (int64_t)gc.n1 * gc.n2;
  cx<T> sm = plane_sum(2 * K0 - 1);
  for (int K = K0; K < K1; ++K) {
    const cx<T> s0 = plane_sum(2 * K), sp = plane_sum(2 * K + 1);
    out[(int64_t)K * cplane + ccol] = w[0] * sm + w[1] * s0 + w[2] * sp;
    sm = sp;
  }
}

// Same full weighting on a real nodal coefficient field (m, gamma).
__global__ __launch_bounds__(256) void k_restrict_real(Grid gf, Grid gc, const double* __restrict__ f,
                                                       double* __restrict__ c) {
  const int64_t pc = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (pc >= gc.N) return;
  const int I = (int)(pc % gc.n1);
  const int64_t q = pc / gc.n1;
  const int J = (int)(q % gc.n2);
  const int K = (int)(q / gc.n2);
  const double w[3] = {0.25, 0.5, 0.25};
  double acc = 0.0;
  const int kk = gf.dim == 3 ? 3 : 1;
  for (int cc = 0; cc < kk; ++cc) {
    int fk = 0;
    double wk = 1.0;
    if (gf.dim == 3) {
      int fz = 2 * K + cc - 1;
      fk = fz < 0 ? 1 : (fz >= gf.n3 ? gf.n3 - 2 : fz);
      wk = w[cc];
    }
    for (int bb = 0; bb < 3; ++bb) {
      int f2 = 2 * J + bb - 1;
      int fj = f2 < 0 ? 1 : (f2 >= gf.n2 ? gf.n2 - 2 : f2);
      for (int a = 0; a < 3; ++a) {
        int f1 = 2 * I + a - 1;
        int fi = f1 < 0 ? 1 : (f1 >= gf.n1 ? gf.n1 - 2 : f1);
        acc += wk * w[bb] * w[a] * f[((int64_t)fk * gf.n2 + fj) * gf.n1 + fi];
      }
    }
  }
  c[pc] = acc;
}

// Bi/tri-linear nodal prolongation of the coarse correction, added in place.
template <typename T>
__global__ __launch_bounds__(256) void k_prolong_add(Grid gf, Grid gc, const cx<T>* __restrict__ xc,
                                                     cx<T>* __restrict__ xf, const int* __restrict__ skip) {
  const int rhs = blockIdx.y;
  if (skip && skip[rhs]) return;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= gf.N) return;
  int i, j, k;
  node_coords(gf, p, i, j, k);
  const cx<T>* c = xc + (int64_t)rhs * gc.N;
  const int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
  const int i1 = (i & 1) ? i0 + 1 : i0, j1 = (j & 1) ? j0 + 1 : j0, k1 = (k & 1) ? k0 + 1 : k0;
  const T wi = (i & 1) ? (T)0.5 : (T)1, wj = (j & 1) ? (T)0.5 : (T)1, wk = (k & 1) ? (T)0.5 : (T)1;
  const int64_t P1 = gc.n1, P2 = (int64_t)gc.n1 * gc.n2;
  cx<T> acc = mk<T>(0, 0);
  const int ik[2] = {i0, i1}, jk[2] = {j0, j1}, kk[2] = {k0, k1};
  const int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1, nk = (k & 1) ? 2 : 1;
  for (int c3 = 0; c3 < nk; ++c3)
    for (int c2 = 0; c2 < nj; ++c2)
      for (int c1 = 0; c1 < ni; ++c1) acc = acc + (wi * wj * wk) * c[kk[c3] * P2 + jk[c2] * P1 + ik[c1]];
  const int64_t o = (int64_t)rhs * gf.N + p;
  xf[o] = xf[o] + acc;
}

// 3-D prolongation marching along z: a thread owns one fine (i, j) column
// over a run of fine planes starting at an even plane.  The in-plane bilinear
// value of each coarse plane K at (i, j) is formed once (weights 1 or 1/2 per
// axis) and reused: fine plane 2K adds c_K, fine plane 2K + 1 adds
// (c_K + c_{K+1}) / 2.
template <typename T>
__global__ __launch_bounds__(256) void k_prolong_add3(Grid gf, Grid gc, int rows, TileMap tm,
                                                      const cx<T>* __restrict__ xc, cx<T>* __restrict__ xf,
                                                      const int* __restrict__ skip) {
  int bx, by, rhs;
  tile_of(tm, bx, by, rhs);
  if (skip && skip[rhs]) return;
  const int n1 = gf.n1, n2 = gf.n2;
  const unsigned item = (unsigned)bx * 256u + threadIdx.x, jj = item / (unsigned)n1;
  const int i = (int)(item - jj * (unsigned)n1), j = (int)jj;
  const int k0 = by * rows;
  if (j >= n2 || k0 >= gf.n3) return;
  const int k1 = min(k0 + rows, gf.n3);
  const cx<T>* c = xc + (int64_t)rhs * gc.N;
  cx<T>* x = xf + (int64_t)rhs * gf.N;
  const int i0 = i >> 1, j0 = j >> 1;
  const int i1 = (i & 1) ? i0 + 1 : i0, j1 = (j & 1) ? j0 + 1 : j0;
  const T wi = (i & 1) ? (T)0.5 : (T)1, wj = (j & 1) ? (T)0.5 : (T)1;
  const int64_t P1 = gc.n1, P2 = (int64_t)gc.n1 * gc.n2;
  const int64_t a00 = j0 * P1 + i0, a01 = j0 * P1 + i1, a10 = j1 * P1 + i0, a11 = j1 * P1 + i1;
  const bool two_i = (i & 1) != 0, two_j = (j & 1) != 0;
  auto plane_val = [&](int K) {
    const cx<T>* cp = c + (int64_t)K * P2;
    cx<T> acc = (wi * wj) * cp[a00];
    if (two_i) acc = acc + (wi * wj) * cp[a01];
    if (two_j) {
      acc = acc + (wi * wj) * cp[a10];
      if (two_i) acc = acc + (wi * wj) * cp[a11];
    }
    return acc;
  };
  const int64_t plane = (int64_t)n1 * n2, col = (int64_t)j * n1 + i;
  cx<T> cK = plane_val(k0 >> 1);
  for (int k = k0; k < k1; ++k) {
    cx<T> e;
    if (k & 1) {
      const cx<T> cK1 = plane_val((k >> 1) + 1);
      e = (T)0.5 * (cK + cK1);
      cK = cK1;
    } else {
      e = cK;
    }
    const int64_t o = (int64_t)k * plane + col;
    x[o] = x[o] + e;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero(cx<T>* __restrict__ a, int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p < n) a[p] = mk<T>(0, 0);
}

template <typename A, typename B>
__global__ __launch_bounds__(256) void k_cast(const cx<A>* __restrict__ a, cx<B>* __restrict__ b, int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p < n) b[p] = cvt<A, B>(a[p]);
}

// Nodal diagonal d = -w^2 m (1 - i g/Re w) + i Re(w) sqrt(m) sum_faces 2/h_d
// (src/GetHelmholtz.jl:41-47, 222-247; Sommerfeld rows always use BC = 2).
// complex diagonal -(w^2) m (1 - i gamma / Re w) + Sommerfeld at node p, in double
__device__ __forceinline__ void diag_at(const Grid& g, const double* __restrict__ m, const double* __restrict__ gam,
                                        double wre, double wim, int neumann_top, int sommerfeld, double h1, double h2,
                                        double h3, int64_t p, double& dr, double& di) {
  const double mv = m[p], gv = gam[p];
  const double w2r = wre * wre - wim * wim, w2i = 2.0 * wre * wim;
  // -(w^2) * m * (1 - i g / wre)
  const double ar = 1.0, ai = -gv / wre;
  dr = -(w2r * ar - w2i * ai) * mv;
  di = -(w2r * ai + w2i * ar) * mv;
  if (sommerfeld) {
    const int i = (int)(p % g.n1);
    const int64_t q = p / g.n1;
    const int j = (int)(q % g.n2);
    const int k = (int)(q / g.n2);
    double s = 0.0;
    if (i == 0) s += 2.0 / h1;
    if (i == g.n1 - 1) s += 2.0 / h1;
    if (g.dim == 2) {
      if (j == 0 && !neumann_top) s += 2.0 / h2;
      if (j == g.n2 - 1) s += 2.0 / h2;
    } else {
      if (j == 0) s += 2.0 / h2;
      if (j == g.n2 - 1) s += 2.0 / h2;
      if (k == 0 && !neumann_top) s += 2.0 / h3;
      if (k == g.n3 - 1) s += 2.0 / h3;
    }
    di += wre * s * sqrt(mv);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_diag(Grid g, const double* __restrict__ m, const double* __restrict__ gam,
                                              double wre, double wim, int neumann_top, int sommerfeld,
                                              double h1, double h2, double h3, cx<T>* __restrict__ d) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.N) return;
  double dr, di;
  diag_at(g, m, gam, wre, wim, neumann_top, sommerfeld, h1, h2, h3, p, dr, di);
  d[p] = mk<T>((T)dr, (T)di);
}

// damped-Jacobi factor relax * (d + Laplacian diagonal)^-1 per node, in double,
// rounded once to T
template <typename T>
__global__ __launch_bounds__(256) void k_jacobi_factor(Grid g, const double* __restrict__ m,
                                                       const double* __restrict__ gam, double wre, double wim,
                                                       int neumann_top, int sommerfeld, double h1, double h2,
                                                       double h3, double relax, cx<T>* __restrict__ f) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.N) return;
  double dr, di;
  diag_at(g, m, gam, wre, wim, neumann_top, sommerfeld, h1, h2, h3, p, dr, di);
  const int i = (int)(p % g.n1);
  const int64_t q = p / g.n1;
  const int j = (int)(q % g.n2);
  const int k = (int)(q / g.n2);
  dr += lap_diag1(i, g.n1, g.bc) * g.ih1 + lap_diag1(j, g.n2, g.bc) * g.ih2;
  if (g.dim == 3) dr += lap_diag1(k, g.n3, g.bc) * g.ih3;
  const double s = relax / (dr * dr + di * di);
  f[p] = mk<T>((T)(dr * s), (T)(-di * s));
}

__global__ __launch_bounds__(256) void k_add_const(double* __restrict__ a, double c, int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p < n) a[p] += c;
}

// ---- coarsest level: dense operator, Gauss-Jordan inverse, dense apply ----
// Column-major n x 2n augmented matrix [A | I]; after elimination the right
// half holds A^-1.
__global__ __launch_bounds__(256) void k_dense_assemble(Grid g, const cx<double>* __restrict__ d,
                                                        cx<double>* __restrict__ M) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n = g.N;
  if (p >= n) return;
  const int i = (int)(p % g.n1);
  const int64_t q = p / g.n1;
  const int j = (int)(q % g.n2);
  const int k = (int)(q / g.n2);
  const int64_t P1 = g.n1, P2 = (int64_t)g.n1 * g.n2;
  // row p of A: sum_d ih_d (2 x_p - x_lo - x_hi) + d_p x_p
  double diag = 2.0 * (g.ih1 + g.ih2 + (g.dim == 3 ? g.ih3 : 0.0));
  M[p + p * n] = mk<double>(diag + d[p].re, d[p].im);
  int64_t nb[6];
  double cf[6];
  int cnt = 0;
  nb[cnt] = (int64_t)k * P2 + j * P1 + lo_nb(i, g.bc); cf[cnt++] = g.ih1;
  nb[cnt] = (int64_t)k * P2 + j * P1 + hi_nb(i, g.n1, g.bc); cf[cnt++] = g.ih1;
  nb[cnt] = (int64_t)k * P2 + lo_nb(j, g.bc) * P1 + i; cf[cnt++] = g.ih2;
  nb[cnt] = (int64_t)k * P2 + hi_nb(j, g.n2, g.bc) * P1 + i; cf[cnt++] = g.ih2;
  if (g.dim == 3) {
    nb[cnt] = (int64_t)lo_nb(k, g.bc) * P2 + j * P1 + i; cf[cnt++] = g.ih3;
    nb[cnt] = (int64_t)hi_nb(k, g.n3, g.bc) * P2 + j * P1 + i; cf[cnt++] = g.ih3;
  }
  for (int c = 0; c < cnt; ++c) {
    cx<double>& e = M[p + nb[c] * n];
    e.re -= cf[c];
  }
  M[p + (n + p) * n] = mk<double>(1.0, 0.0);
}

// pivot search in column k (rows k..n-1): one block of 256 threads; stores
// the pivot row and the reciprocal of the pivot value.
__global__ __launch_bounds__(256) void k_gj_pivot(const cx<double>* __restrict__ M, int64_t n, int64_t k,
                                                  int64_t* __restrict__ piv, cx<double>* __restrict__ pinv) {
  __shared__ double sv[256];
  __shared__ int64_t si[256];
  double best = -1.0;
  int64_t bi = k;
  for (int64_t r = k + threadIdx.x; r < n; r += 256) {
    double v = cabs2(M[r + k * n]);
    if (v > best) { best = v; bi = r; }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double o = sv[threadIdx.x + s];
      int64_t oi = si[threadIdx.x + s];
      if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = o;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    piv[0] = si[0];
    pinv[0] = cinv(M[si[0] + k * n]);
  }
}

// elimination factors of column k as they stand after the row swap k <-> piv
__global__ __launch_bounds__(256) void k_gj_fac(const cx<double>* __restrict__ M, int64_t n, int64_t k,
                                                const int64_t* __restrict__ piv, cx<double>* __restrict__ fac) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const int64_t pr = piv[0];
  cx<double> f = (r == pr) ? M[k + k * n] : M[r + k * n];
  if (r == k) f = mk<double>(0.0, 0.0);
  fac[r] = f;
}

// swap rows k and piv and scale the new row k; one thread per column
__global__ __launch_bounds__(256) void k_gj_swap_scale(cx<double>* __restrict__ M, int64_t n, int64_t k,
                                                       const int64_t* __restrict__ piv,
                                                       const cx<double>* __restrict__ pinv) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= 2 * n) return;
  const int64_t pr = piv[0];
  const cx<double> a = M[k + c * n];
  const cx<double> b = M[pr + c * n];
  if (pr != k) M[pr + c * n] = a;
  M[k + c * n] = b * pinv[0];
}

__global__ __launch_bounds__(256) void k_gj_elim(cx<double>* __restrict__ M, int64_t n, int64_t k,
                                                 const cx<double>* __restrict__ fac) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t c = (int64_t)blockIdx.y + k;  // columns k..2n-1
  if (r >= n || c >= 2 * n || r == k) return;
  cx<double> f = fac[r];
  if (f.re == 0.0 && f.im == 0.0) return;
  M[r + c * n] = M[r + c * n] - f * M[k + c * n];
}

// After elimination the left half of [A | I] is the identity and the right
// half is A^-1 (column-major); copy A^-1 row-major into the left half so a
// row is contiguous for the apply below.
__global__ __launch_bounds__(256) void k_inv_rows(cx<double>* __restrict__ M, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n * n) return;
  const int64_t i = k / n, j = k - i * n;
  M[k] = M[n * n + j * n + i];
}

// x_b = Ainv r_b in complex128 for every right-hand side.  One wavefront per
// RT rows and DR right-hand sides: lanes stride the columns of the row-major
// inverse (coalesced), each lane keeps RT x DR partial sums, wave shuffles
// combine them.  The inverse is re-read nrhs/DR times and r n/RT times.
constexpr int DR = 4, RT = 4;
template <typename T>
__global__ __launch_bounds__(256) void k_dense_apply(const cx<double>* __restrict__ Arow, int64_t n, int nrhs,
                                                     const cx<T>* __restrict__ r, cx<T>* __restrict__ x,
                                                     const int* __restrict__ skip) {
  const int lane = threadIdx.x & 63;
  const int64_t i0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RT;
  const int rhs0 = blockIdx.y * DR;
  if (i0 >= n) return;  // uniform per wavefront
  const int nr = min(DR, nrhs - rhs0);
  const int ni = (int)min((int64_t)RT, n - i0);
  cx<double> acc[RT][DR];
#pragma unroll
  for (int a = 0; a < RT; ++a)
#pragma unroll
    for (int k = 0; k < DR; ++k) acc[a][k] = mk<double>(0.0, 0.0);
  for (int64_t j = lane; j < n; j += 64) {
    cx<double> rv[DR];
#pragma unroll
    for (int k = 0; k < DR; ++k)
      rv[k] = k < nr ? cvt<T, double>(r[(int64_t)(rhs0 + k) * n + j]) : mk<double>(0.0, 0.0);
#pragma unroll
    for (int a = 0; a < RT; ++a) {
      if (a >= ni) break;
      const cx<double> av = Arow[(i0 + a) * n + j];
#pragma unroll
      for (int k = 0; k < DR; ++k) acc[a][k] = acc[a][k] + av * rv[k];
    }
  }
#pragma unroll
  for (int a = 0; a < RT; ++a)
#pragma unroll
    for (int k = 0; k < DR; ++k) {
      double re = acc[a][k].re, im = acc[a][k].im;
      for (int o = 32; o > 0; o >>= 1) {
        re += __shfl_down(re, o, 64);
        im += __shfl_down(im, o, 64);
      }
      if (lane == 0 && a < ni && k < nr && !(skip && skip[rhs0 + k]))
        x[(int64_t)(rhs0 + k) * n + i0 + a] = cvt<double, T>(mk<double>(re, im));
    }
}

// ---- coarsest apply on MFMA (complex64 preconditioner) ----
// The complex product x = A^-1 r is one real GEMM with the real-ified matrix
// A_R (2n x 2n; entry (i, j) of A^-1 becomes [[re, -im], [im, re]] at rows
// 2i, 2i+1 and columns 2j, 2j+1) and r read as a real 2n x nrhs matrix, which
// the interleaved complex64 block already is.  K runs in groups of 16; within
// group g, step u (0..3) of v_mfma_f32_16x16x4_f32 gives lane l the column
// k = 16g + 4(l/16) + u (a permutation of the group, the same for A and r), so
// a lane reads its four A entries as one packed float4 and its four r entries
// as two aligned float2.  A_R is packed once, in fp32: block (R, g), lane l
// holds row 16R + l % 16, columns 16g + 4(l/16) + 0..3 (zero beyond 2n).
__global__ __launch_bounds__(256) void k_pack_inv_mfma(const cx<double>* __restrict__ Arow, int n, int ng, int nrb,
                                                       float4* __restrict__ P) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)nrb * ng * 64) return;
  const int lane = (int)(t & 63);
  const int64_t blk = t >> 6;
  const int g = (int)(blk % ng), R = (int)(blk / ng);
  const int row = R * 16 + (lane & 15), col0 = g * 16 + 4 * (lane >> 4);
  float v[4];
  for (int u = 0; u < 4; ++u) {
    const int col = col0 + u;
    v[u] = 0.f;
    if (row < 2 * n && col < 2 * n) {
      const cx<double> e = Arow[(int64_t)(row >> 1) * n + (col >> 1)];
      const int ra = row & 1, cb = col & 1;
      v[u] = (float)(ra == cb ? e.re : (ra == 0 ? -e.im : e.im));
    }
  }
  P[t] = make_float4(v[0], v[1], v[2], v[3]);
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// one block per 16 x 16 tile of the real-ified result (16 real rows, 16 RHS);
// its four wavefronts take every fourth K-group and their partial tiles are
// summed through LDS in a fixed order (deterministic)
__global__ __launch_bounds__(256) void k_dense_apply_mfma(const float4* __restrict__ P, int n2, int ng, int nrhs,
                                                          const float* __restrict__ r, float* __restrict__ x,
                                                          const int* __restrict__ skip) {
  __shared__ f32x4 part[3][64];
  const int lane = threadIdx.x & 63, wv = (int)(threadIdx.x >> 6);
  const int R = blockIdx.x, C = blockIdx.y;
  int live = 0;
  for (int q = 0; q < 16; ++q) {
    const int cq = C * 16 + q;
    live |= cq < nrhs && !(skip && skip[cq]);
  }
  if (!live) return;  // whole block: every RHS of the tile is done
  const int col = C * 16 + (lane & 15);
  const bool colok = col < nrhs;
  const int kl = lane >> 4;
  const float4* __restrict__ Pb = P + (int64_t)R * ng * 64 + lane;
  // n2 is even, so every column starts 8-byte aligned
  const float2* __restrict__ rc = (const float2*)(r + (int64_t)(colok ? col : 0) * n2);
  f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
  for (int g = wv; g < ng; g += 4) {
    const float4 av = Pb[(int64_t)g * 64];
    const int k = g * 16 + 4 * kl;
    const float2 z2 = make_float2(0.f, 0.f);
    const float2 b01 = (colok && k < n2) ? rc[k >> 1] : z2;
    const float2 b23 = (colok && k + 2 < n2) ? rc[(k >> 1) + 1] : z2;
    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, b01.x, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, b01.y, acc1, 0, 0, 0);
    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, b23.x, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, b23.y, acc1, 0, 0, 0);
  }
  const f32x4 acc = acc0 + acc1;
  if (wv > 0) part[wv - 1][lane] = acc;
  __syncthreads();
  if (wv != 0) return;
  const f32x4 sum = ((acc + part[0][lane]) + part[1][lane]) + part[2][lane];
  if (colok && !(skip && skip[col])) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = R * 16 + 4 * kl + i;
      if (row < n2) x[(int64_t)col * n2 + row] = sum[i];
    }
  }
}

// ---------------------------------------------------------------- launchers
void pack_inv_mfma(const cx<double>* Ainv, int64_t n, float* P, hipStream_t s) {
  const int n2 = (int)(2 * n), ng = (n2 + 15) / 16, nrb = (n2 + 15) / 16;
  const int64_t total = (int64_t)nrb * ng * 64;
  hipLaunchKernelGGL(k_pack_inv_mfma, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, Ainv, (int)n, ng, nrb,
                     (float4*)P);
}
int64_t pack_inv_mfma_size(int64_t n) {
  const int64_t b = (2 * n + 15) / 16;
  return b * b * 64 * 4;
}
void launch_dense_apply_mfma(const float* P, int64_t n, int nrhs, const cx<float>* r, cx<float>* x, const int* skip,
                             hipStream_t s) {
  const int n2 = (int)(2 * n), ng = (n2 + 15) / 16, nrb = (n2 + 15) / 16;
  dim3 grid((unsigned)nrb, (unsigned)((nrhs + 15) / 16));
  hipLaunchKernelGGL(k_dense_apply_mfma, grid, dim3(256), 0, s, (const float4*)P, n2, ng, nrhs, (const float*)r,
                     (float*)x, skip);
}
// blocks: (positions of a row/plane x runs) / 256 along x, runs of planes along y (3-D)
static inline TileMap stencil_map(const Grid& g, int nrhs, int rows) {
  if (g.dim == 2) {
    const int64_t items = (int64_t)g.n1 * ((g.n2 + rows - 1) / rows);
    return make_tile_map((int)((items + 255) / 256), 1, nrhs);
  }
  const int64_t items = (int64_t)g.n1 * g.n2;
  return make_tile_map((int)((items + 255) / 256), (g.n3 + rows - 1) / rows, nrhs);
}

static inline int pick_rows(const Grid& g, int nrhs) {
  // enough independent wavefronts to fill 256 CUs, long runs otherwise
  const int64_t cols = ((int64_t)g.n1 * (g.dim == 3 ? g.n2 : 1) + 63) / 64 * nrhs;
  const int len = g.dim == 3 ? g.n3 : g.n2;
  int rows = 16;
  while (rows > 2 && cols * ((len + rows - 1) / rows) < 8192) rows >>= 1;
  return rows;
}

template <typename TI, typename TO, typename TD, int MODE>
void launch_stencil(const Grid& g, int nrhs, const cx<TI>* x, const cx<TO>* b, cx<TO>* y, const cx<TD>* d,
                    const double* m, cx<double> sfac, double wj, const int* skip, hipStream_t s, int rows_override,
                    const cx<TD>* fj) {
  const int rows = rows_override > 0 ? rows_override : pick_rows(g, nrhs);
  const TileMap tm = stencil_map(g, nrhs, rows);
  if (g.dim == 2)
    hipLaunchKernelGGL((k_stencil2d<TI, TO, TD, MODE>), dim3(tm.T), dim3(256), 0, s, g, x, b, y, d, m, sfac,
                       (TD)wj, rows, tm, skip, fj);
  else
    hipLaunchKernelGGL((k_stencil3d<TI, TO, TD, MODE>), dim3(tm.T), dim3(256), 0, s, g, x, b, y, d, m, sfac,
                       (TD)wj, rows, tm, skip, fj);
}

template void launch_stencil<double, double, double, M_APPLY>(const Grid&, int, const cx<double>*,
                                                               const cx<double>*, cx<double>*,
                                                               const cx<double>*, const double*, cx<double>,
                                                               double, const int*, hipStream_t, int,
                                                              const cx<double>*);
template void launch_stencil<double, double, double, M_RESID>(const Grid&, int, const cx<double>*,
                                                               const cx<double>*, cx<double>*,
                                                               const cx<double>*, const double*, cx<double>,
                                                               double, const int*, hipStream_t, int,
                                                              const cx<double>*);
template void launch_stencil<double, double, double, M_JACOBI>(const Grid&, int, const cx<double>*,
                                                                const cx<double>*, cx<double>*,
                                                                const cx<double>*, const double*, cx<double>,
                                                                double, const int*, hipStream_t, int,
                                                              const cx<double>*);
template void launch_stencil<float, float, float, M_APPLY>(const Grid&, int, const cx<float>*, const cx<float>*,
                                                            cx<float>*, const cx<float>*, const double*,
                                                            cx<double>, double, const int*, hipStream_t, int,
                                                              const cx<float>*);
template void launch_stencil<float, float, float, M_RESID>(const Grid&, int, const cx<float>*, const cx<float>*,
                                                            cx<float>*, const cx<float>*, const double*,
                                                            cx<double>, double, const int*, hipStream_t, int,
                                                              const cx<float>*);
template void launch_stencil<float, float, float, M_JACOBI>(const Grid&, int, const cx<float>*, const cx<float>*,
                                                             cx<float>*, const cx<float>*, const double*,
                                                             cx<double>, double, const int*, hipStream_t, int,
                                                              const cx<float>*);
template void launch_stencil<float, double, double, M_APPLY>(const Grid&, int, const cx<float>*,
                                                              const cx<double>*, cx<double>*,
                                                              const cx<double>*, const double*, cx<double>,
                                                              double, const int*, hipStream_t, int,
                                                              const cx<double>*);

template <typename T>
void launch_jacobi0(const Grid& g, int nrhs, const cx<T>* b, cx<T>* x, const cx<T>* f, const int* skip,
                    hipStream_t s) {
  dim3 grid((unsigned)((g.N + 255) / 256), nrhs);
  hipLaunchKernelGGL((k_jacobi0<T>), grid, dim3(256), 0, s, g, b, x, f, skip);
}
template void launch_jacobi0<double>(const Grid&, int, const cx<double>*, cx<double>*, const cx<double>*,
                                     const int*, hipStream_t);
template void launch_jacobi0<float>(const Grid&, int, const cx<float>*, cx<float>*, const cx<float>*, const int*,
                                    hipStream_t);

template <typename T>
void launch_restrict(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* rf, cx<T>* bc, const int* skip,
                     hipStream_t s) {
  if (gf.dim == 3) {
    const int rows = 4;
    const int64_t items = (int64_t)gc.n1 * gc.n2;
    const TileMap tm = make_tile_map((int)((items + 255) / 256), (gc.n3 + rows - 1) / rows, nrhs);
    hipLaunchKernelGGL((k_restrict3<T>), dim3(tm.T), dim3(256), 0, s, gf, gc, rows, tm, rf, bc, skip);
    return;
  }
  dim3 grid((unsigned)((gc.N + 255) / 256), nrhs);
  hipLaunchKernelGGL((k_restrict<T>), grid, dim3(256), 0, s, gf, gc, rf, bc, skip);
}
template void launch_restrict<double>(const Grid&, const Grid&, int, const cx<double>*, cx<double>*, const int*,
                                      hipStream_t);
template void launch_restrict<float>(const Grid&, const Grid&, int, const cx<float>*, cx<float>*, const int*,
                                     hipStream_t);

void launch_restrict_real(const Grid& gf, const Grid& gc, const double* f, double* c, hipStream_t s) {
  hipLaunchKernelGGL(k_restrict_real, dim3((unsigned)((gc.N + 255) / 256)), dim3(256), 0, s, gf, gc, f, c);
}

template <typename T>
void launch_prolong_add(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* xc, cx<T>* xf, const int* skip,
                        hipStream_t s) {
  if (gf.dim == 3) {
    const int rows = 8;  // even: every run starts on a fine plane that coincides with a coarse one
    const int64_t items = (int64_t)gf.n1 * gf.n2;
    const TileMap tm = make_tile_map((int)((items + 255) / 256), (gf.n3 + rows - 1) / rows, nrhs);
    hipLaunchKernelGGL((k_prolong_add3<T>), dim3(tm.T), dim3(256), 0, s, gf, gc, rows, tm, xc, xf, skip);
    return;
  }
  dim3 grid((unsigned)((gf.N + 255) / 256), nrhs);
  hipLaunchKernelGGL((k_prolong_add<T>), grid, dim3(256), 0, s, gf, gc, xc, xf, skip);
}
template void launch_prolong_add<double>(const Grid&, const Grid&, int, const cx<double>*, cx<double>*,
                                         const int*, hipStream_t);
template void launch_prolong_add<float>(const Grid&, const Grid&, int, const cx<float>*, cx<float>*, const int*,
                                        hipStream_t);

template <typename T>
void launch_zero(cx<T>* a, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_zero<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n);
}
template void launch_zero<double>(cx<double>*, int64_t, hipStream_t);
template void launch_zero<float>(cx<float>*, int64_t, hipStream_t);

template <typename A, typename B>
void launch_cast(const cx<A>* a, cx<B>* b, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_cast<A, B>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n);
}
template void launch_cast<double, float>(const cx<double>*, cx<float>*, int64_t, hipStream_t);
template void launch_cast<float, double>(const cx<float>*, cx<double>*, int64_t, hipStream_t);

template <typename T>
void launch_diag(const Grid& g, const double* m, const double* gam, double wre, double wim, int neumann_top,
                 int sommerfeld, const double h[3], cx<T>* d, hipStream_t s) {
  hipLaunchKernelGGL((k_diag<T>), dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, s, g, m, gam, wre, wim,
                     neumann_top, sommerfeld, h[0], h[1], h[2], d);
}
template void launch_diag<double>(const Grid&, const double*, const double*, double, double, int, int,
                                  const double*, cx<double>*, hipStream_t);
template void launch_diag<float>(const Grid&, const double*, const double*, double, double, int, int,
                                 const double*, cx<float>*, hipStream_t);

template <typename T>
void launch_jacobi_factor(const Grid& g, const double* m, const double* gam, double wre, double wim, int neumann_top,
                          int sommerfeld, const double* h, double relax, cx<T>* f, hipStream_t s) {
  hipLaunchKernelGGL((k_jacobi_factor<T>), dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, s, g, m, gam, wre, wim,
                     neumann_top, sommerfeld, h[0], h[1], h[2], relax, f);
}
template void launch_jacobi_factor<double>(const Grid&, const double*, const double*, double, double, int, int,
                                           const double*, double, cx<double>*, hipStream_t);
template void launch_jacobi_factor<float>(const Grid&, const double*, const double*, double, double, int, int,
                                          const double*, double, cx<float>*, hipStream_t);

void launch_add_const(double* a, double c, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_add_const, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, c, n);
}

void launch_dense_inverse(const Grid& g, const cx<double>* d, cx<double>* M, int64_t* piv, cx<double>* fac,
                          cx<double>* pinv, hipStream_t s) {
  const int64_t n = g.N;
  hipMemsetAsync(M, 0, sizeof(cx<double>) * n * 2 * n, s);
  hipLaunchKernelGGL(k_dense_assemble, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, d, M);
  for (int64_t k = 0; k < n; ++k) {
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(256), 0, s, M, n, k, piv, pinv);
    hipLaunchKernelGGL(k_gj_fac, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, M, n, k, piv, fac);
    hipLaunchKernelGGL(k_gj_swap_scale, dim3((unsigned)((2 * n + 255) / 256)), dim3(256), 0, s, M, n, k, piv,
                       pinv);
    hipLaunchKernelGGL(k_gj_elim, dim3((unsigned)((n + 255) / 256), (unsigned)(2 * n - k)), dim3(256), 0, s, M, n,
                       k, fac);
  }
  hipLaunchKernelGGL(k_inv_rows, dim3((unsigned)((n * n + 255) / 256)), dim3(256), 0, s, M, n);
}

template <typename T>
void launch_dense_apply(const cx<double>* Ainv, int64_t n, int nrhs, const cx<T>* r, cx<T>* x, const int* skip,
                        hipStream_t s) {
  dim3 grid((unsigned)((n + 4 * RT - 1) / (4 * RT)), (unsigned)((nrhs + DR - 1) / DR));
  hipLaunchKernelGGL((k_dense_apply<T>), grid, dim3(256), 0, s, Ainv, n, nrhs, r, x, skip);
}
template void launch_dense_apply<double>(const cx<double>*, int64_t, int, const cx<double>*, cx<double>*,
                                         const int*, hipStream_t);
template void launch_dense_apply<float>(const cx<double>*, int64_t, int, const cx<float>*, cx<float>*,
                                        const int*, hipStream_t);

}  // namespace hh
