// C ABI of the learned coarse-grid correction: network lifecycle, parameter
// access, host-pointer forward / loss-gradient hooks (used by the gradient
// checks), Adam, the training-pair dataset and the training loop.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/helmholtz_amd.h"
#include "hh_unet.h"

using namespace hh;

namespace {
void ck2(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}
template <typename T>
T* dmal2(size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, (n ? n : 1) * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    throw std::runtime_error("hipMalloc failed (" + std::to_string(n * sizeof(T)) + " bytes)");
  }
  return (T*)p;
}

__global__ __launch_bounds__(256) void k_feats_to_x16(const float* __restrict__ f, int64_t npix, bf16* __restrict__ x) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= npix) return;
  for (int c = 0; c < UNET_CIN; ++c) {
    float v = c < UNET_FEATS ? f[k * UNET_FEATS + c] : 0.f;
    unsigned u = __builtin_bit_cast(unsigned, v);
    u = u + 0x7fffu + ((u >> 16) & 1u);
    x[k * UNET_CIN + c] = (bf16)(u >> 16);
  }
}

// training batch: gather samples, normalise by rms(r), scale targets by 1/(s h^2)
__global__ __launch_bounds__(256) void k_batch_rms(const cx<float>* __restrict__ r, int64_t N, const int* __restrict__ idx,
                                                   float* __restrict__ scale) {
  __shared__ float sh[4];
  const cx<float>* rb = r + (int64_t)idx[blockIdx.x] * N;
  float s = 0.f;
  for (int64_t p = threadIdx.x; p < N; p += 256) s += cabs2(rb[p]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float t = sqrtf((sh[0] + sh[1] + sh[2] + sh[3]) / (float)N);
    scale[blockIdx.x] = t > 0.f ? t : 1.f;
  }
}

__global__ __launch_bounds__(256) void k_batch_prep(const cx<float>* __restrict__ r, const cx<float>* __restrict__ e,
                                                    const float* __restrict__ kh2, const float* __restrict__ gw,
                                                    const float* __restrict__ h2, int64_t N, const int* __restrict__ idx,
                                                    const float* __restrict__ scale, int nb, bf16* __restrict__ x,
                                                    float* __restrict__ t) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= N * nb) return;
  const int b = (int)(k / N);
  const int64_t p = k - (int64_t)b * N;
  const int64_t src = (int64_t)idx[b] * N + p;
  const float is = 1.f / scale[b];
  const float it = is / h2[idx[b]];
  const float feats[4] = {r[src].re * is, r[src].im * is, kh2[src], gw[src]};
  for (int c = 0; c < UNET_CIN; ++c) {
    float v = c < 4 ? feats[c] : 0.f;
    unsigned u = __builtin_bit_cast(unsigned, v);
    u = u + 0x7fffu + ((u >> 16) & 1u);
    x[k * UNET_CIN + c] = (bf16)(u >> 16);
  }
  t[2 * k] = e[src].re * it;
  t[2 * k + 1] = e[src].im * it;
}

// relative error of the network's correction per sample, accumulated per block
__global__ __launch_bounds__(256) void k_relerr_parts(const float* __restrict__ y, const float* __restrict__ t, int64_t N,
                                                      float* __restrict__ out) {
  __shared__ float sa[4], sb[4];
  const int b = blockIdx.x;
  float a = 0.f, c = 0.f;
  for (int64_t p = threadIdx.x; p < N; p += 256) {
    const int64_t k = (int64_t)b * N + p;
    const float d0 = y[2 * k] - t[2 * k], d1 = y[2 * k + 1] - t[2 * k + 1];
    a += d0 * d0 + d1 * d1;
    c += t[2 * k] * t[2 * k] + t[2 * k + 1] * t[2 * k + 1];
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    c += __shfl_down(c, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const float num = sa[0] + sa[1] + sa[2] + sa[3], den = sb[0] + sb[1] + sb[2] + sb[3];
    out[b] = den > 0.f ? sqrtf(num / den) : 0.f;
  }
}

thread_local std::string g_learn_err;

template <typename F>
int guard(F&& f) {
  try {
    f();
    return HH_OK;
  } catch (const std::exception& e) {
    g_learn_err = e.what();
    return HH_ERR_HIP;
  }
}
}  // namespace

void hh_set_last_error(const char* msg);  // hh_api.hip

void hh_unet::ensure_stage(int64_t npix) {
  if (npix <= stage_pix) return;
  for (void* p : {(void*)x16, (void*)y, (void*)dy, (void*)tgt, (void*)loss_parts})
    if (p) (void)hipFree(p);
  x16 = dmal2<bf16>((size_t)npix * UNET_CIN);
  y = dmal2<float>((size_t)npix * 2);
  dy = dmal2<float>((size_t)npix * 2);
  tgt = dmal2<float>((size_t)npix * 2);
  loss_parts = dmal2<float>(256);
  stage_pix = npix;
}

hh_unet::~hh_unet() {
  for (void* p : {(void*)x16, (void*)y, (void*)dy, (void*)tgt, (void*)loss_parts})
    if (p) (void)hipFree(p);
  unet_free(net);
  if (stream) (void)hipStreamDestroy(stream);
}

hh_dataset::~hh_dataset() {
  for (void* p : {(void*)r, (void*)e, (void*)kh2, (void*)gw, (void*)h2})
    if (p) (void)hipFree(p);
}

#define LEARN_GUARD(...)                                \
  int rc__ = guard([&] { __VA_ARGS__; });               \
  if (rc__ != HH_OK) hh_set_last_error(g_learn_err.c_str()); \
  return rc__;

extern "C" {

int64_t hh_unet_num_params(void) { return unet_param_count(); }

int hh_unet_create(uint64_t seed, hh_unet** out) {
  LEARN_GUARD({
    if (!out) throw std::runtime_error("null argument");
    hh_unet* u = new hh_unet();
    try {
      ck2(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking), "stream");
      unet_init(u->net, seed, u->stream);
      ck2(hipStreamSynchronize(u->stream), "init");
    } catch (...) {
      delete u;
      throw;
    }
    *out = u;
  })
}

int hh_unet_destroy(hh_unet* u) {
  LEARN_GUARD(delete u)
}

int hh_unet_get_params(const hh_unet* u, float* out) {
  LEARN_GUARD({
    if (!u || !out) throw std::runtime_error("null argument");
    ck2(hipMemcpy(out, u->net.flat, sizeof(float) * u->net.nparam, hipMemcpyDeviceToHost), "get_params");
  })
}

int hh_unet_set_params(hh_unet* u, const float* in) {
  LEARN_GUARD({
    if (!u || !in) throw std::runtime_error("null argument");
    ck2(hipMemcpy(u->net.flat, in, sizeof(float) * u->net.nparam, hipMemcpyHostToDevice), "set_params");
    unet_sync_bf16(u->net);
    ck2(hipStreamSynchronize(u->stream), "set_params");
  })
}

static void check_dims(int nx, int ny) {
  if (nx < 9 || ny < 9 || (nx - 1) % 8 != 0 || (ny - 1) % 8 != 0)
    throw std::runtime_error("U-Net grids need nx, ny = 8k + 1 >= 9 nodes (three nodal halvings)");
}

int hh_unet_forward(hh_unet* u, int nimg, int nx, int ny, const float* feats, float* y) {
  LEARN_GUARD({
    if (!u || !feats || !y || nimg < 1) throw std::runtime_error("bad argument");
    check_dims(nx, ny);
    const int64_t npix = (int64_t)nimg * nx * ny;
    u->ensure_stage(npix);
    float* df = dmal2<float>((size_t)npix * UNET_FEATS);
    ck2(hipMemcpy(df, feats, sizeof(float) * npix * UNET_FEATS, hipMemcpyHostToDevice), "upload");
    hipLaunchKernelGGL(k_feats_to_x16, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, u->stream, df, npix, u->x16);
    unet_forward(u->net, u->ws, u->stream, nimg, nx, ny, u->x16, u->y);
    ck2(hipStreamSynchronize(u->stream), "forward");
    ck2(hipMemcpy(y, u->y, sizeof(float) * npix * 2, hipMemcpyDeviceToHost), "download");
    (void)hipFree(df);
  })
}

int hh_unet_loss_grad(hh_unet* u, int nimg, int nx, int ny, const float* feats, const float* target, float* loss,
                      float* grad) {
  LEARN_GUARD({
    if (!u || !feats || !target || !loss || nimg < 1) throw std::runtime_error("bad argument");
    check_dims(nx, ny);
    const int64_t npix = (int64_t)nimg * nx * ny;
    u->ensure_stage(npix);
    float* df = dmal2<float>((size_t)npix * UNET_FEATS);
    ck2(hipMemcpy(df, feats, sizeof(float) * npix * UNET_FEATS, hipMemcpyHostToDevice), "upload");
    ck2(hipMemcpy(u->tgt, target, sizeof(float) * npix * 2, hipMemcpyHostToDevice), "upload");
    hipLaunchKernelGGL(k_feats_to_x16, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, u->stream, df, npix, u->x16);
    unet_forward(u->net, u->ws, u->stream, nimg, nx, ny, u->x16, u->y);
    unet_mse(npix * 2, u->y, u->tgt, u->dy, u->loss_parts, 64, u->stream);
    unet_backward(u->net, u->ws, u->stream, nimg, nx, ny, u->x16, u->dy);
    ck2(hipStreamSynchronize(u->stream), "backward");
    float parts[64];
    ck2(hipMemcpy(parts, u->loss_parts, sizeof(parts), hipMemcpyDeviceToHost), "loss");
    double s = 0;
    for (float v : parts) s += v;
    *loss = (float)s;
    if (grad) ck2(hipMemcpy(grad, u->net.grad, sizeof(float) * u->net.nparam, hipMemcpyDeviceToHost), "grad");
    (void)hipFree(df);
  })
}

int hh_unet_adam_step(hh_unet* u, const float* grad, float lr, float beta1, float beta2, float eps) {
  LEARN_GUARD({
    if (!u) throw std::runtime_error("null argument");
    if (grad) ck2(hipMemcpy(u->net.grad, grad, sizeof(float) * u->net.nparam, hipMemcpyHostToDevice), "grad");
    unet_adam(u->net, lr, beta1, beta2, eps);
    ck2(hipStreamSynchronize(u->stream), "adam");
  })
}

int hh_dataset_create(int nx, int ny, int64_t capacity, hh_dataset** out) {
  LEARN_GUARD({
    if (!out || capacity < 1) throw std::runtime_error("bad argument");
    check_dims(nx, ny);
    hh_dataset* d = new hh_dataset();
    try {
      const int64_t N = (int64_t)nx * ny;
      d->nx = nx;
      d->ny = ny;
      d->cap = capacity;
      d->r = dmal2<cx<float>>((size_t)capacity * N);
      d->e = dmal2<cx<float>>((size_t)capacity * N);
      d->kh2 = dmal2<float>((size_t)capacity * N);
      d->gw = dmal2<float>((size_t)capacity * N);
      d->h2 = dmal2<float>((size_t)capacity);
    } catch (...) {
      delete d;
      throw;
    }
    *out = d;
  })
}

int hh_dataset_destroy(hh_dataset* d) {
  LEARN_GUARD(delete d)
}

int hh_dataset_size(const hh_dataset* d, int64_t* n) {
  LEARN_GUARD({
    if (!d || !n) throw std::runtime_error("null argument");
    *n = d->n;
  })
}

int hh_unet_train(hh_unet* u, hh_dataset* d, int epochs, int batch, float lr, uint64_t seed, float* loss_per_epoch) {
  LEARN_GUARD({
    if (!u || !d || epochs < 1 || batch < 1) throw std::runtime_error("bad argument");
    if (d->n < 1) throw std::runtime_error("empty dataset");
    const int nb = (int)std::min<int64_t>(batch, d->n);
    const int64_t N = (int64_t)d->nx * d->ny;
    u->ensure_stage(N * nb);
    int* didx = dmal2<int>(nb);
    float* dscale = dmal2<float>(nb);
    std::vector<int> order(d->n);
    std::iota(order.begin(), order.end(), 0);
    std::mt19937_64 rng(seed);
    float parts[64];
    for (int ep = 0; ep < epochs; ++ep) {
      std::shuffle(order.begin(), order.end(), rng);
      double acc = 0;
      int64_t steps = 0;
      for (int64_t b0 = 0; b0 + nb <= d->n; b0 += nb) {
        ck2(hipMemcpyAsync(didx, order.data() + b0, sizeof(int) * nb, hipMemcpyHostToDevice, u->stream), "idx");
        hipLaunchKernelGGL(k_batch_rms, dim3(nb), dim3(256), 0, u->stream, d->r, N, didx, dscale);
        hipLaunchKernelGGL(k_batch_prep, dim3((unsigned)((N * nb + 255) / 256)), dim3(256), 0, u->stream, d->r, d->e,
                           d->kh2, d->gw, d->h2, N, didx, dscale, nb, u->x16, u->tgt);
        unet_forward(u->net, u->ws, u->stream, nb, d->nx, d->ny, u->x16, u->y);
        unet_mse(N * nb * 2, u->y, u->tgt, u->dy, u->loss_parts, 64, u->stream);
        unet_backward(u->net, u->ws, u->stream, nb, d->nx, d->ny, u->x16, u->dy);
        unet_adam(u->net, lr, 0.9f, 0.999f, 1e-8f);
        ck2(hipMemcpyAsync(parts, u->loss_parts, sizeof(parts), hipMemcpyDeviceToHost, u->stream), "loss");
        ck2(hipStreamSynchronize(u->stream), "train step");
        double s = 0;
        for (float v : parts) s += v;
        acc += s;
        ++steps;
      }
      if (loss_per_epoch) loss_per_epoch[ep] = (float)(acc / std::max<int64_t>(1, steps));
    }
    (void)hipFree(didx);
    (void)hipFree(dscale);
    ck2(hipGetLastError(), "train");
  })
}

int hh_unet_eval(hh_unet* u, hh_dataset* d, int64_t first, int64_t count, float* rel_err) {
  LEARN_GUARD({
    if (!u || !d || !rel_err || first < 0 || count < 1 || first + count > d->n) throw std::runtime_error("bad argument");
    const int64_t N = (int64_t)d->nx * d->ny;
    const int nb = (int)std::min<int64_t>(count, 64);
    u->ensure_stage(N * nb);
    int* didx = dmal2<int>(nb);
    float* dscale = dmal2<float>(nb);
    float* derr = dmal2<float>(nb);
    std::vector<int> idx(nb);
    std::vector<float> err(nb);
    double tot = 0;
    for (int64_t b0 = first; b0 < first + count; b0 += nb) {
      const int m = (int)std::min<int64_t>(nb, first + count - b0);
      for (int k = 0; k < m; ++k) idx[k] = (int)(b0 + k);
      ck2(hipMemcpyAsync(didx, idx.data(), sizeof(int) * m, hipMemcpyHostToDevice, u->stream), "idx");
      hipLaunchKernelGGL(k_batch_rms, dim3(m), dim3(256), 0, u->stream, d->r, N, didx, dscale);
      hipLaunchKernelGGL(k_batch_prep, dim3((unsigned)((N * m + 255) / 256)), dim3(256), 0, u->stream, d->r, d->e,
                         d->kh2, d->gw, d->h2, N, didx, dscale, m, u->x16, u->tgt);
      unet_forward(u->net, u->ws, u->stream, m, d->nx, d->ny, u->x16, u->y);
      hipLaunchKernelGGL(k_relerr_parts, dim3(m), dim3(256), 0, u->stream, u->y, u->tgt, N, derr);
      ck2(hipMemcpyAsync(err.data(), derr, sizeof(float) * m, hipMemcpyDeviceToHost, u->stream), "err");
      ck2(hipStreamSynchronize(u->stream), "eval");
      for (int k = 0; k < m; ++k) tot += err[k];
    }
    *rel_err = (float)(tot / count);
    (void)hipFree(didx);
    (void)hipFree(dscale);
    (void)hipFree(derr);
  })
}

}  // extern "C"
