// Learned coarse-grid correction: a 4-level U-Net (16/32/64/64 channels,
// 3x3 convolutions, ReLU) in BF16 with FP32 accumulation on MFMA, with
// hand-written forward, backward and Adam kernels (hh_unet.hip).
#pragma once
#include <vector>

#include "hh_common.h"

namespace hh {

typedef unsigned short bf16;

constexpr int UNET_NCONV = 14;   // 3x3 conv layers
constexpr int UNET_CIN = 16;     // input channels after zero padding (4 used)
constexpr int UNET_FEATS = 4;    // Re r, Im r, k^2 h^2, gamma / Re w

struct UNetLevelDims {
  int nx[4], ny[4];
};

// Activation workspace of one forward/backward context: the network's own
// training hooks, or one solver's inference.  The weights (UNet) are shared
// read-only between contexts; a workspace never is, so solvers on different
// streams that hold one network do not touch each other's activations.
struct UNetWork {
  int nimg = 0, nx = 0, ny = 0;
  std::vector<void*> bufs;
  float* scratch = nullptr;       // wgrad partials
  int64_t scratch_n = 0;
  UNetWork() = default;
  UNetWork(const UNetWork&) = delete;
  UNetWork& operator=(const UNetWork&) = delete;
  void release();
  ~UNetWork() { release(); }
};

struct UNet {
  // parameters (device, fp32 master + bf16 copies)
  float* w[UNET_NCONV] = {};      // [9][co][ci]
  float* b[UNET_NCONV] = {};      // [co]
  bf16* wb[UNET_NCONV] = {};      // forward copy [9][co][ci]
  bf16* wt[UNET_NCONV] = {};      // backward-data copy [9][ci][co], taps flipped
  float* hw = nullptr;            // head 1x1: [2][16]
  float* hb = nullptr;            // [2]
  float* flat = nullptr;          // all parameters, contiguous (w, b per layer, then head)
  float* grad = nullptr;          // same layout
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  int64_t nparam = 0;
  int64_t adam_t = 0;
  hipStream_t stream = nullptr;   // stream of set-up and Adam updates
};

int64_t unet_param_count();
void unet_init(UNet& net, uint64_t seed, hipStream_t s);
void unet_free(UNet& net);
void unet_sync_bf16(UNet& net);
// x: [nimg][nx*ny][16] bf16 input; y: [nimg][nx*ny][2] fp32 output.  Reads
// only the weights; activations go to ws, launches to stream s.
void unet_forward(const UNet& net, UNetWork& ws, hipStream_t s, int nimg, int nx, int ny, const bf16* x, float* y);
// after unet_forward with the same ws and input: dy -> net.grad (overwritten)
void unet_backward(UNet& net, UNetWork& ws, hipStream_t s, int nimg, int nx, int ny, const bf16* x, const float* dy);
void unet_adam(UNet& net, float lr, float b1, float b2, float eps);

// solver-side helpers: complex residual block -> network input, and back
// x16[img][p][16] from r[img][N] (c64) scaled by 1/rms(r_img), feature fields
void unet_prep_input(int nimg, int64_t N, const cx<float>* r, const float* kh2, const float* gw, bf16* x16,
                     float* scale, const int* skip, hipStream_t s);
// e[img][p] = scale[img] * h2 * (y0 + i y1)
void unet_post_output(int nimg, int64_t N, const float* y, const float* scale, float h2, cx<float>* e,
                      const int* skip, hipStream_t s);
// training loss (mean squared error) and its gradient
void unet_mse(int64_t n, const float* y, const float* t, float* dy, float* loss_parts, int nparts, hipStream_t s);

}  // namespace hh

// ---- C-ABI objects (opaque to callers) ----
struct hh_unet {
  hh::UNet net;
  hh::UNetWork ws;  // workspace of this handle's own hooks (training, forward, eval)
  hipStream_t stream = nullptr;
  // staging for the host-pointer hooks and the solver's inference
  hh::bf16* x16 = nullptr;
  float* y = nullptr;
  float* dy = nullptr;
  float* tgt = nullptr;
  float* loss_parts = nullptr;
  int64_t stage_pix = 0;
  void ensure_stage(int64_t npix);
  ~hh_unet();
};

struct hh_dataset {
  int nx = 0, ny = 0;
  int64_t cap = 0, n = 0;
  hh::cx<float>* r = nullptr;  // [cap][N] residuals (network input)
  hh::cx<float>* e = nullptr;  // [cap][N] errors (targets)
  float* kh2 = nullptr;        // [cap][N]
  float* gw = nullptr;         // [cap][N]
  float* h2 = nullptr;         // [cap]
  ~hh_dataset();
};
