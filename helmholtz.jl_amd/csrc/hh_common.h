// Shared device-side definitions for the Helmholtz SL-MG kernels (gfx950).
//
// Layout in HBM: every field is nodal and column-major (x fastest), one
// contiguous block of N = n1*n2[*n3] interleaved complex values per
// right-hand side, RHS blocks back to back: value (p, b) lives at p + b*N.
// This is exactly Julia's ComplexF64 N x nrhs matrix, so the boundary needs no
// transpose.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hh {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> mk(T r, T i) {
  cx<T> z;
  z.re = r;
  z.im = i;
  return z;
}
template <typename A, typename T>
__host__ __device__ __forceinline__ cx<T> cvt(const cx<A>& a) {
  return mk<T>((T)a.re, (T)a.im);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) {
  return mk<T>(a.re + b.re, a.im + b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) {
  return mk<T>(a.re - b.re, a.im - b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return mk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(T s, cx<T> a) {
  return mk<T>(s * a.re, s * a.im);
}
// a * b with explicit fused multiply-adds: the plain operator leaves the
// compiler free to contract either product of each component into an fma,
// and its choice can differ between two instantiations of one kernel
template <typename T>
__host__ __device__ __forceinline__ cx<T> cmul_fma(cx<T> a, cx<T> b) {
  return mk<T>(fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re));
}
// conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
  return mk<T>(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> cinv(cx<T> a) {
  T d = T(1) / (a.re * a.re + a.im * a.im);
  return mk<T>(a.re * d, -a.im * d);
}
template <typename T>
__host__ __device__ __forceinline__ T cabs2(cx<T> a) {
  return a.re * a.re + a.im * a.im;
}

// Geometry of one grid level as the kernels see it.
struct Grid {
  int dim;          // 2 or 3
  int n1, n2, n3;   // nodes per axis (n3 = 1 in 2-D)
  int64_t N;        // nodes
  double ih1, ih2, ih3;  // 1/h^2 per axis
  int bc;           // 2: mirror ghost (second-order Neumann rows), 1: ghost = self
};

// Node coordinates of p (i fastest), in 32-bit arithmetic when the grid has
// fewer than 2^31 nodes.
__device__ __forceinline__ void node_coords(const Grid& g, int64_t p, int& i, int& j, int& k) {
  if (g.N < ((int64_t)1 << 31)) {
    const unsigned u = (unsigned)p, q = u / (unsigned)g.n1;
    i = (int)(u - q * (unsigned)g.n1);
    j = (int)(q % (unsigned)g.n2);
    k = (int)(q / (unsigned)g.n2);
  } else {
    const int64_t q = p / g.n1;
    i = (int)(p - q * g.n1);
    j = (int)(q % g.n2);
    k = (int)(q / g.n2);
  }
}

// mirror ghost index for tiled kernels; positions farther out than one
// mirror (only loaded by a last tile, never used for an owned output) are
// clamped into the grid
__device__ __forceinline__ int mirror_idx(int g, int n) {
  g = g < 0 ? -g : (g >= n ? 2 * (n - 1) - g : g);
  return g < 0 ? 0 : (g >= n ? n - 1 : g);
}

// Mirror/self ghost index on one axis (src/PlainNodalLaplacian.jl:18-30 rows).
__device__ __forceinline__ int lo_nb(int i, int bc) { return i > 0 ? i - 1 : (bc == 2 ? 1 : 0); }
__device__ __forceinline__ int hi_nb(int i, int n, int bc) {
  return i < n - 1 ? i + 1 : (bc == 2 ? n - 2 : n - 1);
}
// Diagonal of the 1-D second difference at index i (times h^2).
__device__ __forceinline__ double lap_diag1(int i, int n, int bc) {
  return (bc == 1 && (i == 0 || i == n - 1)) ? 1.0 : 2.0;
}

// A x at one node of the 5-point (2-D) or 7-point (3-D) operator from the
// node value xc, its neighbours along each axis and the diagonal dv:
//   sum_d ih_d (2 xc - x_{-d} - x_{+d}) + dv xc
// with a fixed operation order (explicit fused multiply-adds), so every
// kernel that forms it -- the stencil launches, the fused multigrid legs and
// the Arnoldi passes that form H z on the fly -- gets the same bits.
template <typename T>
__device__ __forceinline__ T ax_dxc_re(cx<T> dv, cx<T> xc) {
  return fma(dv.re, xc.re, -(dv.im * xc.im));
}
template <typename T>
__device__ __forceinline__ T ax_dxc_im(cx<T> dv, cx<T> xc) {
  return fma(dv.re, xc.im, dv.im * xc.re);
}
template <typename T>
__device__ __forceinline__ cx<T> apply5(cx<T> xc, cx<T> xl, cx<T> xr, cx<T> xm, cx<T> xp, T ih1, T ih2, cx<T> dv) {
  const T a1r = ((T)2 * xc.re - xl.re) - xr.re, a1i = ((T)2 * xc.im - xl.im) - xr.im;
  const T a2r = ((T)2 * xc.re - xm.re) - xp.re, a2i = ((T)2 * xc.im - xm.im) - xp.im;
  return mk<T>(fma(ih1, a1r, ih2 * a2r) + ax_dxc_re(dv, xc), fma(ih1, a1i, ih2 * a2i) + ax_dxc_im(dv, xc));
}
template <typename T>
__device__ __forceinline__ cx<T> apply7(cx<T> xc, cx<T> x1l, cx<T> x1r, cx<T> x2l, cx<T> x2r, cx<T> x3l,
                                        cx<T> x3r, T ih1, T ih2, T ih3, cx<T> dv) {
  const T a1r = ((T)2 * xc.re - x1l.re) - x1r.re, a1i = ((T)2 * xc.im - x1l.im) - x1r.im;
  const T a2r = ((T)2 * xc.re - x2l.re) - x2r.re, a2i = ((T)2 * xc.im - x2l.im) - x2r.im;
  const T a3r = ((T)2 * xc.re - x3l.re) - x3r.re, a3i = ((T)2 * xc.im - x3l.im) - x3r.im;
  return mk<T>(fma(ih1, a1r, fma(ih2, a2r, ih3 * a3r)) + ax_dxc_re(dv, xc),
               fma(ih1, a1i, fma(ih2, a2i, ih3 * a3i)) + ax_dxc_im(dv, xc));
}

// XCD-aware work order.  A launch is a 1-D grid of nxt * nyt * nrhs blocks;
// the hardware hands block b to XCD b % 8, and each XCD has its own L2.  Block
// b is given the (b / 8)-th item of a contiguous 1/8 slice of the work list
// ordered (RHS group, tile row, RHS within the group, tile column), so each
// XCD walks its own RHS slab row by row: a tile row's coefficients are reused
// by the G RHS of the group and the halo rows of the previous tile row (same
// RHS) are still in that XCD's L2.
struct TileMap {
  int nxt, nyt, G;
  unsigned T;  // nxt * nyt * nrhs < 2^31
};

inline TileMap make_tile_map(int nxt, int nyt, int nrhs) {
  TileMap m;
  m.nxt = nxt;
  m.nyt = nyt;
  m.G = nrhs % 4 == 0 ? 4 : (nrhs % 2 == 0 ? 2 : 1);
  m.T = (unsigned)((int64_t)nxt * nyt * nrhs);
  return m;
}

__device__ __forceinline__ void tile_of(const TileMap& m, int& xt, int& yt, int& rhs) {
  const unsigned b = blockIdx.x;
  const unsigned per = m.T >> 3, rem = m.T & 7u;
  const unsigned xcd = b & 7u;
  unsigned L = xcd * per + (xcd < rem ? xcd : rem) + (b >> 3);
  xt = (int)(L % (unsigned)m.nxt);
  L /= (unsigned)m.nxt;
  const int r_in = (int)(L % (unsigned)m.G);
  L /= (unsigned)m.G;
  yt = (int)(L % (unsigned)m.nyt);
  rhs = (int)(L / (unsigned)m.nyt) * m.G + r_in;
}

}  // namespace hh
