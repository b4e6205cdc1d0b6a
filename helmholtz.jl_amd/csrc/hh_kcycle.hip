// K-cycle coarse correction (Notay & Vassilevski's recursive Krylov cycle):
// two steps of a minimal-residual (GCR) iteration on the coarse level,
// preconditioned by the coarse cycle, batched over right-hand sides:
//   c1 = B rc, v1 = A c1, a1 = <v1, rc>/<v1, v1>, r2 = rc - a1 v1,
//   c2 = B r2, v2 = A c2, g = <v1, v2>/<v1, v1>, v2 -= g v1, c2 -= g c1,
//   a2 = <v2, r2>/<v2, v2>, e = a1 c1 + a2 c2.
// The per-RHS inner products are reduced in a fixed order (deterministic);
// scalars are complex128, vectors stay in the level's precision.
#include "hh_kernels.h"

namespace hh {

namespace {

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// block (256) sum, result in thread 0
__device__ __forceinline__ double bsum(double v, double* sh) {
  v = wsum(v);
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

}  // namespace

// part[(c * 2 + k) * nrhs + b] = sum over chunk c of conj(a_k) b_k, k < nd
template <typename T>
__global__ __launch_bounds__(256) void k_kc_dots(int64_t N, int nrhs, int nch, const cx<T>* __restrict__ a0,
                                                 const cx<T>* __restrict__ b0, const cx<T>* __restrict__ a1,
                                                 const cx<T>* __restrict__ b1, int nd, cx<double>* __restrict__ part,
                                                 const int* __restrict__ skip) {
  const int c = blockIdx.x, b = blockIdx.y;
  if (skip && skip[b]) return;
  __shared__ double sh[4];
  const int64_t L = (N + nch - 1) / nch, p0 = (int64_t)c * L, p1 = min(N, p0 + L);
  const int64_t base = (int64_t)b * N;
  cx<double> s0 = mk<double>(0.0, 0.0), s1 = mk<double>(0.0, 0.0);
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    s0 = s0 + cmulc(up(a0[base + p]), up(b0[base + p]));
    if (nd > 1) s1 = s1 + cmulc(up(a1[base + p]), up(b1[base + p]));
  }
  for (int k = 0; k < nd; ++k) {
    const cx<double> v = k == 0 ? s0 : s1;
    const double re = bsum(v.re, sh), im = bsum(v.im, sh);
    if (threadIdx.x == 0) part[((int64_t)c * 2 + k) * nrhs + b] = mk<double>(re, im);
  }
}

// per-RHS scalars sc[k * nrhs + b]: k = 0 a1, 1 <v1, v1>, 2 g, 3 a2.
// stage 1: a1 from (<v1, rc>, <v1, v1>) ; stage 2: g from <v1, v2> ;
// stage 3: a2 from (<v2, r2>, <v2, v2>).  One wavefront per RHS.
__global__ __launch_bounds__(64) void k_kc_scalar(int nrhs, int nch, const cx<double>* __restrict__ part, int stage,
                                                  cx<double>* __restrict__ sc, const int* __restrict__ skip) {
  const int b = blockIdx.x;
  if (skip && skip[b]) return;
  double r0 = 0.0, i0 = 0.0, r1 = 0.0, i1 = 0.0;
  for (int c = threadIdx.x; c < nch; c += 64) {
    const cx<double> x0 = part[((int64_t)c * 2 + 0) * nrhs + b];
    r0 += x0.re;
    i0 += x0.im;
    if (stage != 2) {
      const cx<double> x1 = part[((int64_t)c * 2 + 1) * nrhs + b];
      r1 += x1.re;
      i1 += x1.im;
    }
  }
  r0 = wsum(r0);
  i0 = wsum(i0);
  r1 = wsum(r1);
  i1 = wsum(i1);
  if (threadIdx.x != 0) return;
  const cx<double> d0 = mk<double>(r0, i0);
  if (stage == 1) {  // d0 = <v1, rc>, r1 = <v1, v1>
    sc[1 * nrhs + b] = mk<double>(r1, 0.0);
    sc[0 * nrhs + b] = r1 > 0.0 ? mk<double>(d0.re / r1, d0.im / r1) : mk<double>(0.0, 0.0);
  } else if (stage == 2) {  // d0 = <v1, v2>
    const double n1 = sc[1 * nrhs + b].re;
    sc[2 * nrhs + b] = n1 > 0.0 ? mk<double>(d0.re / n1, d0.im / n1) : mk<double>(0.0, 0.0);
  } else {  // d0 = <v2, r2>, r1 = <v2, v2>
    sc[3 * nrhs + b] = r1 > 0.0 ? mk<double>(d0.re / r1, d0.im / r1) : mk<double>(0.0, 0.0);
  }
}

// mode 0: r2 = rc - a1 v1 (in place on y = rc, x = v1)
// mode 1: v2 -= g v1 (y = v2, x = v1) and c2 -= g c1 (y2 = c2, x2 = c1)
// mode 2: e = a1 c1 + a2 c2 (y = c2 in place, x = c1)
template <typename T>
__global__ __launch_bounds__(256) void k_kc_update(int64_t N, int nrhs, int mode, const cx<double>* __restrict__ sc,
                                                   cx<T>* __restrict__ y, const cx<T>* __restrict__ x,
                                                   cx<T>* __restrict__ y2, const cx<T>* __restrict__ x2,
                                                   const int* __restrict__ skip) {
  const int b = blockIdx.y;
  if (skip && skip[b]) return;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  const int64_t o = (int64_t)b * N + p;
  if (mode == 0) {
    const cx<double> a1 = sc[0 * nrhs + b];
    const cx<double> v = up(y[o]) - a1 * up(x[o]);
    y[o] = mk<T>((T)v.re, (T)v.im);
  } else if (mode == 1) {
    const cx<double> g = sc[2 * nrhs + b];
    const cx<double> v = up(y[o]) - g * up(x[o]);
    const cx<double> c = up(y2[o]) - g * up(x2[o]);
    y[o] = mk<T>((T)v.re, (T)v.im);
    y2[o] = mk<T>((T)c.re, (T)c.im);
  } else {
    const cx<double> a1 = sc[0 * nrhs + b], a2 = sc[3 * nrhs + b];
    const cx<double> e = a1 * up(x[o]) + a2 * up(y[o]);
    y[o] = mk<T>((T)e.re, (T)e.im);
  }
}

int kc_chunks(int64_t N) { return (int)std::max<int64_t>(1, std::min<int64_t>(64, (N + 4095) / 4096)); }

template <typename T>
void kc_dots(int64_t N, int nrhs, const cx<T>* a0, const cx<T>* b0, const cx<T>* a1, const cx<T>* b1, int nd,
             cx<double>* part, const int* skip, hipStream_t s) {
  hipLaunchKernelGGL((k_kc_dots<T>), dim3(kc_chunks(N), nrhs), dim3(256), 0, s, N, nrhs, kc_chunks(N), a0, b0, a1, b1,
                     nd, part, skip);
}
void kc_scalar(int64_t N, int nrhs, const cx<double>* part, int stage, cx<double>* sc, const int* skip,
               hipStream_t s) {
  hipLaunchKernelGGL(k_kc_scalar, dim3(nrhs), dim3(64), 0, s, nrhs, kc_chunks(N), part, stage, sc, skip);
}
template <typename T>
void kc_update(int64_t N, int nrhs, int mode, const cx<double>* sc, cx<T>* y, const cx<T>* x, cx<T>* y2,
               const cx<T>* x2, const int* skip, hipStream_t s) {
  hipLaunchKernelGGL((k_kc_update<T>), dim3((unsigned)((N + 255) / 256), nrhs), dim3(256), 0, s, N, nrhs, mode, sc,
                     y, x, y2, x2, skip);
}

#define HH_INST(T)                                                                                            \
  template void kc_dots<T>(int64_t, int, const cx<T>*, const cx<T>*, const cx<T>*, const cx<T>*, int,         \
                           cx<double>*, const int*, hipStream_t);                                             \
  template void kc_update<T>(int64_t, int, int, const cx<double>*, cx<T>*, const cx<T>*, cx<T>*, const cx<T>*, \
                             const int*, hipStream_t);
HH_INST(float)
HH_INST(double)
#undef HH_INST

}  // namespace hh
