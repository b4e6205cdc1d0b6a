// U-Net coarse-grid correction: BF16 activations/weights with FP32
// accumulation on MFMA (v_mfma_f32_16x16x16_bf16), hand-written forward,
// backward and Adam.
//
// Layout: activations NHWC with the pixel index x-fastest (p = i + nx*j, the
// solver's nodal order) and channels contiguous (multiples of 16).
// 3x3 convolution = implicit GEMM: out[p][co] = sum_{tap,ci} W[tap][co][ci]
// in[p + off(tap)][ci] (zero padding).  One wavefront computes 16 pixels x all
// output channels (the A operand, 16 pixels x 16 input channels, is loaded once
// and reused for every 16-channel output block).  Backward-data is the same
// kernel on the output gradient with W transposed and taps flipped; the
// weight gradient is a split-K MFMA over pixels with per-chunk partials summed
// in a fixed order (deterministic).  Down-sampling is nodal full weighting
// ([1 2 1]/4 per axis, stride 2, zero padding) and up-sampling nodal
// bilinear interpolation; their adjoints are each other up to a factor 4.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "hh_unet.h"

namespace hh {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __host__ __forceinline__ bf16 f2bf(float f) {
  unsigned u = __builtin_bit_cast(unsigned, f);
  u = u + 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __host__ __forceinline__ float bf2f(bf16 h) {
  unsigned u = ((unsigned)h) << 16;
  return __builtin_bit_cast(float, u);
}

struct ConvSpec {
  int ci, co, level;
};
// encoder 16/32/64, bottleneck 64, decoder with skip concatenations
static const ConvSpec kSpec[UNET_NCONV] = {{16, 16, 0},  {16, 16, 0},  {16, 32, 1}, {32, 32, 1}, {32, 64, 2},
                                           {64, 64, 2},  {64, 64, 3},  {64, 64, 3}, {128, 64, 2}, {64, 64, 2},
                                           {96, 32, 1},  {32, 32, 1},  {48, 16, 0}, {16, 16, 0}};

int64_t unet_param_count() {
  int64_t n = 0;
  for (int l = 0; l < UNET_NCONV; ++l) n += 9 * kSpec[l].co * kSpec[l].ci + kSpec[l].co;
  return n + 2 * 16 + 2;
}

static void ck(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}
template <typename T>
static T* dmal(size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, (n ? n : 1) * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    throw std::runtime_error("hipMalloc failed in U-Net workspace");
  }
  return (T*)p;
}

// ------------------------------------------------------------------ kernels
// A wavefront computes PM x 16 pixels x NCO * 16 output channels (channel
// block blockIdx.z of a layer with `cout` outputs).  Every operand load of a
// (tap, 16-channel) step feeds several MFMAs: the input tile of a pixel group
// is used for the NCO channel blocks, the weight tile for the PM pixel groups
// (with one of each, two 512-byte loads per MFMA kept the loads, not the
// matrix cores, busy).  NCO and PM are picked per launch so that small grids
// still give several wavefronts per SIMD.
template <int NCO, int PM>
__global__ __launch_bounds__(256) void k_conv3x3(const bf16* __restrict__ in, int cin, const bf16* __restrict__ w,
                                                 const float* __restrict__ bias, bf16* __restrict__ out, int nx,
                                                 int ny, int relu, int cout) {
  const int img = blockIdx.y;
  const int npix = nx * ny;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tile = (blockIdx.x * 4 + wv) * 16 * PM;
  if (tile >= npix) return;  // uniform per wavefront
  const int co0 = blockIdx.z * NCO * 16;
  const int m = lane & 15;
  const int kq = (lane >> 4) * 4;
  int ia[PM], ja[PM];
  bool pa_ok[PM];
#pragma unroll
  for (int g = 0; g < PM; ++g) {
    const int pa = tile + 16 * g + m;
    pa_ok[g] = pa < npix;
    ia[g] = pa_ok[g] ? pa % nx : 0;
    ja[g] = pa_ok[g] ? pa / nx : 0;
  }
  const bf16* inb = in + (int64_t)img * npix * cin;
  f4 acc[PM][NCO];
#pragma unroll
  for (int g = 0; g < PM; ++g)
#pragma unroll
    for (int c = 0; c < NCO; ++c) acc[g][c] = f4{0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < 9; ++tap) {
    const bf16* src[PM];
    bool ok[PM];
#pragma unroll
    for (int g = 0; g < PM; ++g) {
      const int ii = ia[g] + tap % 3 - 1, jj = ja[g] + tap / 3 - 1;
      ok[g] = pa_ok[g] && ii >= 0 && ii < nx && jj >= 0 && jj < ny;
      src[g] = inb + ((int64_t)(ok[g] ? jj : 0) * nx + (ok[g] ? ii : 0)) * cin + kq;
    }
    const bf16* wtap = w + (int64_t)tap * cout * cin + (int64_t)(co0 + m) * cin + kq;
    for (int cb = 0; cb < cin; cb += 16) {
      s4 a[PM];
#pragma unroll
      for (int g = 0; g < PM; ++g) a[g] = ok[g] ? *(const s4*)(src[g] + cb) : s4{0, 0, 0, 0};
#pragma unroll
      for (int c = 0; c < NCO; ++c) {
        const s4 bv = *(const s4*)(wtap + (int64_t)c * 16 * cin + cb);
#pragma unroll
        for (int g = 0; g < PM; ++g) acc[g][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[g], bv, acc[g][c], 0, 0, 0);
      }
    }
  }
  bf16* ob = out + (int64_t)img * npix * cout;
#pragma unroll
  for (int g = 0; g < PM; ++g)
#pragma unroll
    for (int c = 0; c < NCO; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int p = tile + 16 * g + kq + i;
        const int co = co0 + c * 16 + m;
        if (p < npix) {
          float v = acc[g][c][i] + (bias ? bias[co] : 0.f);
          if (relu) v = v > 0.f ? v : 0.f;
          ob[(int64_t)p * cout + co] = f2bf(v);
        }
      }
}

// Weight and bias gradients of a 3x3 layer as one MFMA GEMM per tap over
// pixels: dW[tap][co][ci] = sum_q dpre[q][co] in[q + off(tap)][ci] and
// db[co] = sum_q dpre[q][co].  A block takes the pixel range [q0, q0 + per) of
// the flattened (image, pixel) index and one 16-channel block of ci (grid.y),
// for all 9 taps and all output channels.  Each chunk of 64 pixels is staged
// once in LDS with 16-byte global loads, transposed so that the MFMA's
// reduction index (pixels) is contiguous: sA[co][px] = dpre, and the
// im2col of the input sB[tap][ci][px] (zero outside the image); a thread
// loads 4 pixels x 8 channels and stores each channel's 4 pixels as one
// 8-byte piece (a quarter of the LDS stores of one 2-byte store per value).  Every
// operand of v_mfma_f32_16x16x16bf16 (lane: 4 consecutive pixels of one
// channel) is then one aligned 8-byte LDS read; rows are padded to 68
// entries so the 16 lanes of a group hit distinct banks.  The 9 NCO tiles of
// 16x16 outputs are spread over the 4 wavefronts; partial sums per block are
// reduced in a fixed order afterwards (deterministic).
constexpr int WG_PX = 64;         // pixels per staged chunk
constexpr int WG_LD = WG_PX + 4;  // padded LDS row (136 B: 8-byte aligned pieces)

// Four pixels' 8 channels (v[i]: pixel i, channels 0..7, two per word) into
// the transposed tile: channel k's four pixels as one 8-byte store at
// dst + k WG_LD (dst 8-byte aligned)
__device__ __forceinline__ void store_px4(bf16* dst, const uint4 (&v)[4]) {
  const unsigned w[4][4] = {{v[0].x, v[0].y, v[0].z, v[0].w}, {v[1].x, v[1].y, v[1].z, v[1].w},
                            {v[2].x, v[2].y, v[2].z, v[2].w}, {v[3].x, v[3].y, v[3].z, v[3].w}};
#pragma unroll
  for (int h = 0; h < 4; ++h) {  // word h: channels 2h (low half) and 2h + 1 (high half)
    uint2 lo, hi;
    lo.x = (w[0][h] & 0xffffu) | (w[1][h] << 16);
    lo.y = (w[2][h] & 0xffffu) | (w[3][h] << 16);
    hi.x = (w[0][h] >> 16) | (w[1][h] & 0xffff0000u);
    hi.y = (w[2][h] >> 16) | (w[3][h] & 0xffff0000u);
    *(uint2*)(dst + (2 * h) * WG_LD) = lo;
    *(uint2*)(dst + (2 * h + 1) * WG_LD) = hi;
  }
}
template <int NCO>
__global__ __launch_bounds__(256) void k_conv_wgrad(const bf16* __restrict__ dpre, const bf16* __restrict__ in,
                                                   int cin, int nx, int ny, int64_t total, int64_t per,
                                                   float* __restrict__ part) {
  constexpr int COUT = 16 * NCO;
  constexpr int NTILE = 9 * NCO;
  constexpr int TPW = (NTILE + 3) / 4;
  __shared__ __align__(16) bf16 sA[COUT * WG_LD];
  __shared__ __align__(16) bf16 sB[9 * 16 * WG_LD];
  __shared__ int sbase[WG_PX], si[WG_PX], sj[WG_PX];
  const int cib = blockIdx.y;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, kq = (lane >> 4) * 4;
  const int npix = nx * ny;
  const int64_t qb0 = (int64_t)blockIdx.x * per;
  const int64_t qb1 = qb0 + per < total ? qb0 + per : total;
  f4 acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (int64_t q0 = qb0; q0 < qb1; q0 += WG_PX) {
    __syncthreads();  // the previous chunk's operands are consumed
    if (threadIdx.x < WG_PX) {
      const int64_t q = q0 + threadIdx.x;
      if (q < qb1) {
        const int64_t img = q / npix;
        const int p = (int)(q - img * npix);
        sbase[threadIdx.x] = (int)img;
        si[threadIdx.x] = p % nx;
        sj[threadIdx.x] = p / nx;
      } else {
        sbase[threadIdx.x] = -1;
      }
    }
    // dpre: 64 pixels x COUT channels; an item is 4 pixels x 8 channels (four
    // 16-byte loads), stored as one 8-byte row piece per channel
    for (int e = threadIdx.x; e < (WG_PX / 4) * (COUT / 8); e += 256) {
      const int px4 = e / (COUT / 8), c8 = (e % (COUT / 8)) * 8;
      uint4 v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t q = q0 + 4 * px4 + i;
        v[i] = q < qb1 ? *(const uint4*)(dpre + q * COUT + c8) : make_uint4(0u, 0u, 0u, 0u);
      }
      store_px4(sA + c8 * WG_LD + 4 * px4, v);
    }
    __syncthreads();  // pixel coordinates
    // im2col of the input for this ci block: tap t, 4 pixels x 8 channels per item
    for (int e = threadIdx.x; e < 9 * (WG_PX / 4) * 2; e += 256) {
      const int t = e / ((WG_PX / 4) * 2), r = e % ((WG_PX / 4) * 2);
      const int px4 = r >> 1, c8 = (r & 1) * 8;
      uint4 v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int px = 4 * px4 + i;
        v[i] = make_uint4(0u, 0u, 0u, 0u);
        const int img = sbase[px];
        if (img >= 0) {
          const int ii = si[px] + t % 3 - 1, jj = sj[px] + t / 3 - 1;
          if (ii >= 0 && ii < nx && jj >= 0 && jj < ny)
            v[i] = *(const uint4*)(in + ((int64_t)img * npix + (int64_t)jj * nx + ii) * cin + cib * 16 + c8);
        }
      }
      store_px4(sB + (t * 16 + c8) * WG_LD + 4 * px4, v);
    }
    __syncthreads();
    // bias gradient: one thread per output channel (ci block 0 only)
    if (cib == 0 && threadIdx.x < COUT)
      for (int px = 0; px < WG_PX; ++px) bsum += bf2f(sA[threadIdx.x * WG_LD + px]);
#pragma unroll
    for (int k0 = 0; k0 < WG_PX; k0 += 16) {
#pragma unroll
      for (int ti = 0; ti < TPW; ++ti) {
        const int tile = wv + 4 * ti;
        if (tile < NTILE) {
          const int t = tile / NCO, cob = tile % NCO;
          const s4 a = *(const s4*)(sA + (cob * 16 + m) * WG_LD + k0 + kq);
          const s4 b = *(const s4*)(sB + (t * 16 + m) * WG_LD + k0 + kq);
          acc[ti] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[ti], 0, 0, 0);
        }
      }
    }
  }
  // partials part[block][9 COUT cin + COUT]: the weights [tap][co][ci] (C row =
  // co, column = ci), then the bias [co] — the layer's own parameter layout,
  // so one reduction sums both
  const int64_t nw = 9LL * COUT * cin;
  float* out = part + (int64_t)blockIdx.x * (nw + COUT);
#pragma unroll
  for (int ti = 0; ti < TPW; ++ti) {
    const int tile = wv + 4 * ti;
    if (tile < NTILE) {
      const int t = tile / NCO, cob = tile % NCO;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        out[((int64_t)t * COUT + cob * 16 + kq + i) * cin + cib * 16 + m] = acc[ti][i];
    }
  }
  if (cib == 0 && threadIdx.x < COUT) out[nw + threadIdx.x] = bsum;
}

// out[k] = sum_c part[c][k] in a fixed order: a block takes 16 outputs, its
// 16 thread rows sum the partials c = row, row + 16, ... and the first row adds
// the 16 row sums in order (one launch of n / 16 blocks instead of n / 256
// threads each walking all partials: 3-5 us instead of ~20 us per layer)
__global__ __launch_bounds__(256) void k_sum_parts(const float* __restrict__ part, int nparts, int64_t n,
                                                   float* __restrict__ out) {
  __shared__ float sh[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int64_t k = (int64_t)blockIdx.x * 16 + tx;
  float acc = 0.f;
  if (k < n)
    for (int c = ty; c < nparts; c += 16) acc += part[(int64_t)c * n + k];
  sh[ty][tx] = acc;
  __syncthreads();
  if (ty == 0 && k < n) {
    float s = 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) s += sh[r][tx];
    out[k] = s;
  }
}

__global__ __launch_bounds__(256) void k_relu_bwd(const bf16* __restrict__ dout, const bf16* __restrict__ out,
                                                  bf16* __restrict__ dpre, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < n) dpre[k] = bf2f(out[k]) > 0.f ? dout[k] : (bf16)0;
}

// nodal full weighting, stride 2, zero padding; out = scale * R in
__global__ __launch_bounds__(256) void k_down(const bf16* __restrict__ in, int C, int nx, int ny, bf16* __restrict__ out,
                                              int ncx, int ncy, int nimg, float scale) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nc = (int64_t)ncx * ncy;
  if (k >= nc * nimg * C) return;
  const int c = (int)(k % C);
  const int64_t q = k / C;
  const int64_t img = q / nc;
  const int pc = (int)(q - img * nc);
  const int I = pc % ncx, J = pc / ncx;
  const float wgt[3] = {0.25f, 0.5f, 0.25f};
  float s = 0.f;
  for (int b = -1; b <= 1; ++b) {
    const int j = 2 * J + b;
    if (j < 0 || j >= ny) continue;
    for (int a = -1; a <= 1; ++a) {
      const int i = 2 * I + a;
      if (i < 0 || i >= nx) continue;
      s += wgt[a + 1] * wgt[b + 1] * bf2f(in[((img * ny + j) * (int64_t)nx + i) * C + c]);
    }
  }
  out[k] = f2bf(scale * s);
}

// nodal bilinear prolongation; out (+)= scale * P in
__global__ __launch_bounds__(256) void k_up(const bf16* __restrict__ in, int C, int ncx, int ncy, bf16* __restrict__ out,
                                            int nx, int ny, int nimg, float scale, int accumulate) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nf = (int64_t)nx * ny;
  if (k >= nf * nimg * C) return;
  const int c = (int)(k % C);
  const int64_t q = k / C;
  const int64_t img = q / nf;
  const int p = (int)(q - img * nf);
  const int i = p % nx, j = p / nx;
  const int i0 = i >> 1, j0 = j >> 1;
  const int i1 = (i & 1) ? i0 + 1 : i0, j1 = (j & 1) ? j0 + 1 : j0;
  const float wi = (i & 1) ? 0.5f : 1.f, wj = (j & 1) ? 0.5f : 1.f;
  auto at = [&](int a, int b) {
    return (a < ncx && b < ncy) ? bf2f(in[((img * ncy + b) * (int64_t)ncx + a) * C + c]) : 0.f;
  };
  float s;
  if ((i & 1) && (j & 1))
    s = 0.25f * (at(i0, j0) + at(i1, j0) + at(i0, j1) + at(i1, j1));
  else if (i & 1)
    s = 0.5f * (at(i0, j0) + at(i1, j0));
  else if (j & 1)
    s = 0.5f * (at(i0, j0) + at(i0, j1));
  else
    s = at(i0, j0);
  (void)wi;
  (void)wj;
  s *= scale;
  if (accumulate) s += bf2f(out[k]);
  out[k] = f2bf(s);
}

__global__ __launch_bounds__(256) void k_concat(const bf16* __restrict__ a, int Ca, const bf16* __restrict__ b, int Cb,
                                                bf16* __restrict__ out, int64_t npix) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int C = Ca + Cb;
  if (k >= npix * C) return;
  const int c = (int)(k % C);
  const int64_t p = k / C;
  out[k] = c < Ca ? a[p * Ca + c] : b[p * Cb + (c - Ca)];
}

// da = dc[:, :Ca] ; db += dc[:, Ca:]
__global__ __launch_bounds__(256) void k_split(const bf16* __restrict__ dc, int Ca, int Cb, bf16* __restrict__ da,
                                               bf16* __restrict__ db, int accumulate_b, int64_t npix) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int C = Ca + Cb;
  if (k >= npix * C) return;
  const int c = (int)(k % C);
  const int64_t p = k / C;
  if (c < Ca) {
    da[p * Ca + c] = dc[k];
  } else {
    const int64_t o = p * Cb + (c - Ca);
    float v = bf2f(dc[k]);
    if (accumulate_b) v += bf2f(db[o]);
    db[o] = f2bf(v);
  }
}

// 1x1 head: y[p][o] = hb[o] + sum_c hw[o][c] x[p][c], 16 -> 2, fp32 out
__global__ __launch_bounds__(256) void k_head_fwd(const bf16* __restrict__ x, const float* __restrict__ hw,
                                                  const float* __restrict__ hb, float* __restrict__ y, int64_t npix) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= npix) return;
  float y0 = hb[0], y1 = hb[1];
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const float v = bf2f(x[p * 16 + c]);
    y0 += hw[c] * v;
    y1 += hw[16 + c] * v;
  }
  y[2 * p] = y0;
  y[2 * p + 1] = y1;
}

// dx = hw^T dy (bf16); per-block partial sums of dhw (2x16) and dhb (2)
__global__ __launch_bounds__(256) void k_head_bwd(const bf16* __restrict__ x, const float* __restrict__ hw,
                                                  const float* __restrict__ dy, bf16* __restrict__ dx, int64_t npix,
                                                  int64_t chunk, float* __restrict__ part) {
  __shared__ float sh[34][64];
  const int64_t q0 = (int64_t)blockIdx.x * chunk;
  const int64_t q1 = q0 + chunk < npix ? q0 + chunk : npix;
  float acc[34];
#pragma unroll
  for (int k = 0; k < 34; ++k) acc[k] = 0.f;
  for (int64_t p = q0 + threadIdx.x; p < q1; p += 256) {
    const float d0 = dy[2 * p], d1 = dy[2 * p + 1];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const float v = bf2f(x[p * 16 + c]);
      acc[c] += d0 * v;
      acc[16 + c] += d1 * v;
      dx[p * 16 + c] = f2bf(hw[c] * d0 + hw[16 + c] * d1);
    }
    acc[32] += d0;
    acc[33] += d1;
  }
  // reduce 256 threads -> 1 (fixed order): wave shuffles then LDS
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 34; ++k) {
    float v = acc[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) sh[k][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 34) {
    const int k = threadIdx.x;
    part[(int64_t)blockIdx.x * 34 + k] = sh[k][0] + sh[k][1] + sh[k][2] + sh[k][3];
  }
}

__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ t,
                                             float* __restrict__ dy, float* __restrict__ parts, int64_t n, float inv_n) {
  __shared__ float sh[4];
  float s = 0.f;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) {
    const float d = y[k] - t[k];
    dy[k] = 2.f * d * inv_n;
    s += d * d;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) parts[blockIdx.x] = (sh[0] + sh[1] + sh[2] + sh[3]) * inv_n;
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const float gk = g[k];
  const float mk = b1 * m[k] + (1.f - b1) * gk;
  const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  w[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// wb[tap][co][ci] = bf16(w[tap][co][ci]); wt[tap][ci][co] = bf16(w[8 - tap][co][ci])
__global__ __launch_bounds__(256) void k_wcopy(const float* __restrict__ w, bf16* __restrict__ wb, bf16* __restrict__ wt,
                                               int co, int ci) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n = 9LL * co * ci;
  if (k >= n) return;
  const int c_i = (int)(k % ci);
  const int64_t r = k / ci;
  const int c_o = (int)(r % co);
  const int tap = (int)(r / co);
  const bf16 v = f2bf(w[k]);
  wb[k] = v;
  wt[((int64_t)(8 - tap) * ci + c_i) * co + c_o] = v;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}
// normal(0, std) from a counter-based hash (Box-Muller)
__global__ __launch_bounds__(256) void k_randn(float* __restrict__ w, int64_t n, float stdv, uint64_t seed) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const uint64_t h1 = mix64(seed * 0x9e3779b97f4a7c15ULL + 2 * (uint64_t)k + 1);
  const uint64_t h2 = mix64(h1 ^ 0x632be59bd9b4e019ULL);
  const float u1 = ((h1 >> 40) + 1.f) / 16777217.f, u2 = (h2 >> 40) / 16777216.f;
  w[k] = stdv * sqrtf(-2.f * logf(u1)) * cosf(6.28318530718f * u2);
}

// network input: [Re r, Im r] / rms(r), k^2 h^2, gamma / Re w, zeros
__global__ __launch_bounds__(256) void k_rms(const cx<float>* __restrict__ r, int64_t N, float* __restrict__ scale,
                                             const int* __restrict__ skip) {
  __shared__ float sh[4];
  const int img = blockIdx.x;
  if (skip && skip[img]) return;
  const cx<float>* rb = r + (int64_t)img * N;
  float s = 0.f;
  for (int64_t p = threadIdx.x; p < N; p += 256) s += cabs2(rb[p]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float t = sqrtf((sh[0] + sh[1] + sh[2] + sh[3]) / (float)N);
    scale[img] = t > 0.f ? t : 1.f;
  }
}

__global__ __launch_bounds__(256) void k_prep(const cx<float>* __restrict__ r, int64_t N, const float* __restrict__ kh2,
                                              const float* __restrict__ gw, const float* __restrict__ scale,
                                              bf16* __restrict__ x, const int* __restrict__ skip, int nimg) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= N * nimg) return;
  const int64_t img = k / N;
  const int64_t p = k - img * N;
  bf16* xo = x + k * UNET_CIN;
  if (skip && skip[img]) {
    for (int c = 0; c < UNET_CIN; ++c) xo[c] = 0;
    return;
  }
  const float is = 1.f / scale[img];
  const cx<float> v = r[k];
  xo[0] = f2bf(v.re * is);
  xo[1] = f2bf(v.im * is);
  xo[2] = f2bf(kh2[p]);
  xo[3] = f2bf(gw[p]);
#pragma unroll
  for (int c = 4; c < UNET_CIN; ++c) xo[c] = 0;
}

__global__ __launch_bounds__(256) void k_post(const float* __restrict__ y, int64_t N, const float* __restrict__ scale,
                                              float h2, cx<float>* __restrict__ e, const int* __restrict__ skip,
                                              int nimg) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= N * nimg) return;
  const int64_t img = k / N;
  if (skip && skip[img]) return;
  const float s = scale[img] * h2;
  e[k] = mk<float>(s * y[2 * k], s * y[2 * k + 1]);
}

// ------------------------------------------------------------------ host side
static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

enum Act {
  A_A0, A_S0, A_D1, A_A1, A_S1, A_D2, A_A2, A_S2, A_D3, A_A3, A_B3, A_U2, A_C2, A_A4, A_B4, A_U1, A_C1, A_A5, A_B5,
  A_U0, A_C0, A_A6, A_B6, A_NUM
};
static const int kActLevel[A_NUM] = {0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 2, 2, 1, 1, 1, 1, 0, 0, 0, 0};
static const int kActC[A_NUM] = {16, 16, 16, 32, 32, 32, 64, 64, 64, 64, 64, 64, 128, 64, 64, 64, 96, 32, 32, 32, 48, 16, 16};
// conv layer l: input activation, output activation (-1 input = network input)
static const int kIn[UNET_NCONV] = {-1, A_A0, A_D1, A_A1, A_D2, A_A2, A_D3, A_A3, A_C2, A_A4, A_C1, A_A5, A_C0, A_A6};
static const int kOut[UNET_NCONV] = {A_A0, A_S0, A_A1, A_S1, A_A2, A_S2, A_A3, A_B3, A_A4, A_B4, A_A5, A_B5, A_A6, A_B6};

struct Dims {
  int nx[4], ny[4];
  int64_t T[4];
};
static Dims dims_of(int nimg, int nx, int ny) {
  Dims d;
  d.nx[0] = nx;
  d.ny[0] = ny;
  for (int l = 1; l < 4; ++l) {
    d.nx[l] = (d.nx[l - 1] + 1) / 2;
    d.ny[l] = (d.ny[l - 1] + 1) / 2;
  }
  for (int l = 0; l < 4; ++l) d.T[l] = (int64_t)nimg * d.nx[l] * d.ny[l];
  return d;
}

void UNetWork::release() {
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  bufs.clear();
  if (scratch) (void)hipFree(scratch);
  scratch = nullptr;
  scratch_n = 0;
  nimg = nx = ny = 0;
}

// buffers: [0, A_NUM) activations, [A_NUM, 2 A_NUM) gradients, then dpre and
// d(head input).  Reallocated when the shape grows; a smaller batch of the same
// grid reuses them (launches index with their own nimg).
static void ensure_ws(UNetWork& ws, int nimg, int nx, int ny) {
  if (ws.nx == nx && ws.ny == ny && nimg <= ws.nimg && !ws.bufs.empty()) return;
  for (void* p : ws.bufs)
    if (p) (void)hipFree(p);
  ws.bufs.assign(2 * A_NUM + 2, nullptr);
  ws.nimg = ws.nx = ws.ny = 0;
  Dims d = dims_of(nimg, nx, ny);
  for (int a = 0; a < A_NUM; ++a) {
    ws.bufs[a] = dmal<bf16>((size_t)d.T[kActLevel[a]] * kActC[a]);
    ws.bufs[A_NUM + a] = dmal<bf16>((size_t)d.T[kActLevel[a]] * kActC[a]);
  }
  ws.bufs[2 * A_NUM] = dmal<bf16>((size_t)d.T[0] * 64);      // dpre scratch
  ws.bufs[2 * A_NUM + 1] = dmal<bf16>((size_t)d.T[0] * 16);  // d(head input)
  ws.nimg = nimg;
  ws.nx = nx;
  ws.ny = ny;
}

static bf16* act(UNetWork& ws, int a) { return (bf16*)ws.bufs[a]; }
static bf16* grd(UNetWork& ws, int a) { return (bf16*)ws.bufs[A_NUM + a]; }

// output channels (NCO 16-channel blocks) and pixel groups (PM groups of 16)
// per wavefront: the fewest operand loads per MFMA (1 / NCO + 1 / PM) with
// NCO * PM <= 8 (at most 32 accumulator registers: six or more wavefronts per
// SIMD) that still leaves 2048 wavefronts (two per SIMD)
static void conv_shape(int cout, int64_t pixels, int* nco_out, int* pm_out) {
  const int nb = cout / 16;
  const int64_t groups = (pixels + 15) / 16;
  int best_nco = 1, best_pm = 1;
  double best = 2.0;
  for (int nco : {8, 6, 4, 3, 2, 1}) {
    if (nb % nco) continue;
    for (int pm : {4, 2, 1}) {
      if (nco * pm > 8) continue;
      const int64_t waves = (groups + pm - 1) / pm * (nb / nco);
      const double loads = 1.0 / nco + 1.0 / pm;
      if (waves >= 2048 && loads < best - 1e-9) {
        best = loads;
        best_nco = nco;
        best_pm = pm;
      }
    }
  }
  *nco_out = best_nco;
  *pm_out = best_pm;
}

template <int NCO>
static void conv_launch_pm(int pm, dim3 gridbase, int npix, const bf16* in, int cin, const bf16* w,
                           const float* bias, bf16* out, int cout, int nx, int ny, int relu, hipStream_t s) {
  dim3 grid((unsigned)((npix + 64 * pm - 1) / (64 * pm)), gridbase.y, gridbase.z);
  switch (pm) {
    case 1: hipLaunchKernelGGL((k_conv3x3<NCO, 1>), grid, dim3(256), 0, s, in, cin, w, bias, out, nx, ny, relu, cout); break;
    case 2: hipLaunchKernelGGL((k_conv3x3<NCO, 2>), grid, dim3(256), 0, s, in, cin, w, bias, out, nx, ny, relu, cout); break;
    case 4: hipLaunchKernelGGL((k_conv3x3<NCO, 4>), grid, dim3(256), 0, s, in, cin, w, bias, out, nx, ny, relu, cout); break;
    default: throw std::runtime_error("unsupported pixel grouping");
  }
}

static void conv_launch(const bf16* in, int cin, const bf16* w, const float* bias, bf16* out, int cout, int nimg, int nx,
                        int ny, int relu, hipStream_t s) {
  if (cout % 16 || cin % 16 || cin <= 0) throw std::runtime_error("unsupported channel count");
  int nco, pm;
  conv_shape(cout, (int64_t)nimg * nx * ny, &nco, &pm);
  const dim3 gb(1, nimg, (unsigned)(cout / 16 / nco));
  const int npix = nx * ny;
  switch (nco) {
    case 1: conv_launch_pm<1>(pm, gb, npix, in, cin, w, bias, out, cout, nx, ny, relu, s); break;
    case 2: conv_launch_pm<2>(pm, gb, npix, in, cin, w, bias, out, cout, nx, ny, relu, s); break;
    case 3: conv_launch_pm<3>(pm, gb, npix, in, cin, w, bias, out, cout, nx, ny, relu, s); break;
    case 4: conv_launch_pm<4>(pm, gb, npix, in, cin, w, bias, out, cout, nx, ny, relu, s); break;
    case 6: conv_launch_pm<6>(pm, gb, npix, in, cin, w, bias, out, cout, nx, ny, relu, s); break;
    case 8: conv_launch_pm<8>(pm, gb, npix, in, cin, w, bias, out, cout, nx, ny, relu, s); break;
    default: throw std::runtime_error("unsupported channel count");
  }
}

void unet_sync_bf16(UNet& net) {
  for (int l = 0; l < UNET_NCONV; ++l) {
    const int64_t n = 9LL * kSpec[l].co * kSpec[l].ci;
    hipLaunchKernelGGL(k_wcopy, dim3(nblk(n)), dim3(256), 0, net.stream, net.w[l], net.wb[l], net.wt[l], kSpec[l].co,
                       kSpec[l].ci);
  }
}

void unet_init(UNet& net, uint64_t seed, hipStream_t s) {
  net.stream = s;
  net.nparam = unet_param_count();
  net.flat = dmal<float>(net.nparam);
  net.grad = dmal<float>(net.nparam);
  net.adam_m = dmal<float>(net.nparam);
  net.adam_v = dmal<float>(net.nparam);
  ck(hipMemsetAsync(net.flat, 0, sizeof(float) * net.nparam, s), "memset");
  ck(hipMemsetAsync(net.grad, 0, sizeof(float) * net.nparam, s), "memset");
  ck(hipMemsetAsync(net.adam_m, 0, sizeof(float) * net.nparam, s), "memset");
  ck(hipMemsetAsync(net.adam_v, 0, sizeof(float) * net.nparam, s), "memset");
  int64_t off = 0;
  for (int l = 0; l < UNET_NCONV; ++l) {
    const int64_t nw = 9LL * kSpec[l].co * kSpec[l].ci;
    net.w[l] = net.flat + off;
    off += nw;
    net.b[l] = net.flat + off;
    off += kSpec[l].co;
    const int fan_in = 9 * (l == 0 ? UNET_FEATS : kSpec[l].ci);
    hipLaunchKernelGGL(k_randn, dim3(nblk(nw)), dim3(256), 0, s, net.w[l], nw, sqrtf(2.f / fan_in),
                       seed * 131 + l);
    net.wb[l] = dmal<bf16>(nw);
    net.wt[l] = dmal<bf16>(nw);
  }
  net.hw = net.flat + off;
  off += 32;
  net.hb = net.flat + off;
  hipLaunchKernelGGL(k_randn, dim3(1), dim3(256), 0, s, net.hw, (int64_t)32, 0.25f, seed * 131 + 99);
  net.adam_t = 0;
  unet_sync_bf16(net);
  ck(hipGetLastError(), "unet_init");
}

void unet_free(UNet& net) {
  for (int l = 0; l < UNET_NCONV; ++l) {
    if (net.wb[l]) (void)hipFree(net.wb[l]);
    if (net.wt[l]) (void)hipFree(net.wt[l]);
  }
  for (float* p : {net.flat, net.grad, net.adam_m, net.adam_v})
    if (p) (void)hipFree(p);
  net = UNet{};
}

void unet_forward(const UNet& net, UNetWork& ws, hipStream_t s, int nimg, int nx, int ny, const bf16* x, float* y) {
  ensure_ws(ws, nimg, nx, ny);
  Dims d = dims_of(nimg, nx, ny);
  auto conv = [&](int l) {
    const ConvSpec& c = kSpec[l];
    const int lv = c.level;
    const bf16* in = kIn[l] < 0 ? x : act(ws, kIn[l]);
    conv_launch(in, c.ci, net.wb[l], net.b[l], act(ws, kOut[l]), c.co, nimg, d.nx[lv], d.ny[lv], 1, s);
  };
  auto down = [&](int a_in, int a_out, int lf) {
    const int C = kActC[a_in];
    hipLaunchKernelGGL(k_down, dim3(nblk(d.T[lf + 1] * C)), dim3(256), 0, s, act(ws, a_in), C, d.nx[lf], d.ny[lf],
                       act(ws, a_out), d.nx[lf + 1], d.ny[lf + 1], nimg, 1.f);
  };
  auto up = [&](int a_in, int a_out, int lc) {
    const int C = kActC[a_in];
    hipLaunchKernelGGL(k_up, dim3(nblk(d.T[lc - 1] * C)), dim3(256), 0, s, act(ws, a_in), C, d.nx[lc], d.ny[lc],
                       act(ws, a_out), d.nx[lc - 1], d.ny[lc - 1], nimg, 1.f, 0);
  };
  auto cat = [&](int a, int b, int c, int lv) {
    hipLaunchKernelGGL(k_concat, dim3(nblk(d.T[lv] * kActC[c])), dim3(256), 0, s, act(ws, a), kActC[a], act(ws, b),
                       kActC[b], act(ws, c), d.T[lv]);
  };
  conv(0); conv(1); down(A_S0, A_D1, 0);
  conv(2); conv(3); down(A_S1, A_D2, 1);
  conv(4); conv(5); down(A_S2, A_D3, 2);
  conv(6); conv(7);
  up(A_B3, A_U2, 3); cat(A_U2, A_S2, A_C2, 2);
  conv(8); conv(9);
  up(A_B4, A_U1, 2); cat(A_U1, A_S1, A_C1, 1);
  conv(10); conv(11);
  up(A_B5, A_U0, 1); cat(A_U0, A_S0, A_C0, 0);
  conv(12); conv(13);
  hipLaunchKernelGGL(k_head_fwd, dim3(nblk(d.T[0])), dim3(256), 0, s, act(ws, A_B6), net.hw, net.hb, y, d.T[0]);
}

static int64_t pick_chunks(int64_t total, int64_t* chunk) {
  int64_t nch = total / 4096;
  if (nch < 1) nch = 1;
  if (nch > 64) nch = 64;
  int64_t c = (total + nch - 1) / nch;
  c = (c + 15) / 16 * 16;
  *chunk = c;
  return (total + c - 1) / c;
}

// Pixel blocks of the weight-gradient GEMM: a block walks its chunks of WG_PX
// pixels one after another (stage, barrier, MFMA), so few long blocks leave
// the kernel bound by that chain (about 60 us per launch whatever the layer's
// size when every block took at least 16 chunks).  Blocks of WG_MINCH chunks,
// at most WG_MAXBLK of them (past that the partials' reduction costs more
// than the shorter chains save) and as many as the partials' scratch holds for
// the layer (np weights and biases per block).
constexpr int64_t WG_MINCH = 2;
constexpr int64_t WG_MAXBLK = 1024;
constexpr int64_t WG_SCRATCH = 64LL << 20;  // floats of weight-gradient partials
static int64_t pick_wgrad_blocks(int64_t total, int64_t np, int64_t* per) {
  int64_t nb = (total + WG_MINCH * WG_PX - 1) / (WG_MINCH * WG_PX);
  if (nb > WG_MAXBLK) nb = WG_MAXBLK;
  const int64_t cap = WG_SCRATCH / np;
  if (nb > cap) nb = cap;
  if (nb < 1) nb = 1;
  int64_t p = (total + nb - 1) / nb;
  p = (p + WG_PX - 1) / WG_PX * WG_PX;
  *per = p;
  return (total + p - 1) / p;
}

static void ensure_scratch(UNetWork& ws, int64_t n) {
  if (ws.scratch_n >= n) return;
  if (ws.scratch) (void)hipFree(ws.scratch);
  ws.scratch = nullptr;
  ws.scratch_n = 0;
  ws.scratch = dmal<float>(n);
  ws.scratch_n = n;
}

void unet_backward(UNet& net, UNetWork& ws, hipStream_t s, int nimg, int nx, int ny, const bf16* x, const float* dy) {
  Dims d = dims_of(nimg, nx, ny);
  bf16* dpre = (bf16*)ws.bufs[2 * A_NUM];
  // scratch: weight- and bias-gradient partials (WG_SCRATCH floats), or head
  // partials
  ensure_scratch(ws, WG_SCRATCH);
  const int64_t head_off = unet_param_count() - 34;
  // head
  {
    int64_t chunk;
    const int64_t nch = pick_chunks(d.T[0], &chunk);
    hipLaunchKernelGGL(k_head_bwd, dim3((unsigned)nch), dim3(256), 0, s, act(ws, A_B6), net.hw, dy, grd(ws, A_B6),
                       d.T[0], chunk, ws.scratch);
    hipLaunchKernelGGL(k_sum_parts, dim3(3), dim3(256), 0, s, ws.scratch, (int)nch, (int64_t)34, net.grad + head_off);
  }
  int64_t off[UNET_NCONV];
  {
    int64_t o = 0;
    for (int l = 0; l < UNET_NCONV; ++l) {
      off[l] = o;
      o += 9LL * kSpec[l].co * kSpec[l].ci + kSpec[l].co;
    }
  }
  // conv layer backward: g(out) -> dpre -> dW, db, g(in)
  auto conv_b = [&](int l, bool need_din) {
    const ConvSpec& c = kSpec[l];
    const int lv = c.level;
    const int64_t T = d.T[lv];
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(T * c.co)), dim3(256), 0, s, grd(ws, kOut[l]), act(ws, kOut[l]), dpre,
                       T * c.co);
    const bf16* in = kIn[l] < 0 ? x : act(ws, kIn[l]);
    int64_t per;
    const int64_t np = 9LL * c.co * c.ci + c.co;  // the layer's weights and bias
    const int64_t nb = pick_wgrad_blocks(T, np, &per);
    float* pw = ws.scratch;
    const dim3 g((unsigned)nb, (unsigned)(c.ci / 16));
    switch (c.co / 16) {
      case 1: hipLaunchKernelGGL(k_conv_wgrad<1>, g, dim3(256), 0, s, dpre, in, c.ci, d.nx[lv], d.ny[lv], T, per, pw); break;
      case 2: hipLaunchKernelGGL(k_conv_wgrad<2>, g, dim3(256), 0, s, dpre, in, c.ci, d.nx[lv], d.ny[lv], T, per, pw); break;
      case 4: hipLaunchKernelGGL(k_conv_wgrad<4>, g, dim3(256), 0, s, dpre, in, c.ci, d.nx[lv], d.ny[lv], T, per, pw); break;
      default: throw std::runtime_error("unsupported output channel count");
    }
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((np + 15) / 16)), dim3(256), 0, s, pw, (int)nb, np,
                       net.grad + off[l]);
    if (need_din)
      conv_launch(dpre, c.co, net.wt[l], nullptr, grd(ws, kIn[l]), c.ci, nimg, d.nx[lv], d.ny[lv], 0, s);
  };
  auto split = [&](int ac, int au, int askip, int lv) {
    hipLaunchKernelGGL(k_split, dim3(nblk(d.T[lv] * kActC[ac])), dim3(256), 0, s, grd(ws, ac), kActC[au],
                       kActC[askip], grd(ws, au), grd(ws, askip), 0, d.T[lv]);
  };
  // g(b) = P^T g(u) = 4 R g(u)
  auto up_b = [&](int au, int ab, int lc) {
    const int C = kActC[au];
    hipLaunchKernelGGL(k_down, dim3(nblk(d.T[lc] * C)), dim3(256), 0, s, grd(ws, au), C, d.nx[lc - 1], d.ny[lc - 1],
                       grd(ws, ab), d.nx[lc], d.ny[lc], nimg, 4.f);
  };
  // g(s) += R^T g(d) = P g(d) / 4
  auto down_b = [&](int ad, int as, int lf) {
    const int C = kActC[ad];
    hipLaunchKernelGGL(k_up, dim3(nblk(d.T[lf] * C)), dim3(256), 0, s, grd(ws, ad), C, d.nx[lf + 1], d.ny[lf + 1],
                       grd(ws, as), d.nx[lf], d.ny[lf], nimg, 0.25f, 1);
  };
  conv_b(13, true); conv_b(12, true);
  split(A_C0, A_U0, A_S0, 0);
  up_b(A_U0, A_B5, 1);
  conv_b(11, true); conv_b(10, true);
  split(A_C1, A_U1, A_S1, 1);
  up_b(A_U1, A_B4, 2);
  conv_b(9, true); conv_b(8, true);
  split(A_C2, A_U2, A_S2, 2);
  up_b(A_U2, A_B3, 3);
  conv_b(7, true); conv_b(6, true);
  down_b(A_D3, A_S2, 2);
  conv_b(5, true); conv_b(4, true);
  down_b(A_D2, A_S1, 1);
  conv_b(3, true); conv_b(2, true);
  down_b(A_D1, A_S0, 0);
  conv_b(1, true); conv_b(0, false);
}

void unet_adam(UNet& net, float lr, float b1, float b2, float eps) {
  net.adam_t += 1;
  const float c1 = 1.f - powf(b1, (float)net.adam_t), c2 = 1.f - powf(b2, (float)net.adam_t);
  hipLaunchKernelGGL(k_adam, dim3(nblk(net.nparam)), dim3(256), 0, net.stream, net.flat, net.grad, net.adam_m,
                     net.adam_v, net.nparam, lr, b1, b2, eps, c1, c2);
  unet_sync_bf16(net);
}

void unet_prep_input(int nimg, int64_t N, const cx<float>* r, const float* kh2, const float* gw, bf16* x16,
                     float* scale, const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(k_rms, dim3(nimg), dim3(256), 0, s, r, N, scale, skip);
  hipLaunchKernelGGL(k_prep, dim3(nblk(N * nimg)), dim3(256), 0, s, r, N, kh2, gw, scale, x16, skip, nimg);
}

void unet_post_output(int nimg, int64_t N, const float* y, const float* scale, float h2, cx<float>* e, const int* skip,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_post, dim3(nblk(N * nimg)), dim3(256), 0, s, y, N, scale, h2, e, skip, nimg);
}

void unet_mse(int64_t n, const float* y, const float* t, float* dy, float* loss_parts, int nparts, hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(nparts), dim3(256), 0, s, y, t, dy, loss_parts, n, 1.f / (float)n);
}

}  // namespace hh
