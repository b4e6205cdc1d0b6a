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

// Mirror/self ghost index on one axis (src/PlainNodalLaplacian.jl:18-30 rows).
__device__ __forceinline__ int lo_nb(int i, int bc) { return i > 0 ? i - 1 : (bc == 2 ? 1 : 0); }
__device__ __forceinline__ int hi_nb(int i, int n, int bc) {
  return i < n - 1 ? i + 1 : (bc == 2 ? n - 2 : n - 1);
}
// Diagonal of the 1-D second difference at index i (times h^2).
__device__ __forceinline__ double lap_diag1(int i, int n, int bc) {
  return (bc == 1 && (i == 0 || i == n - 1)) ? 1.0 : 2.0;
}

}  // namespace hh
