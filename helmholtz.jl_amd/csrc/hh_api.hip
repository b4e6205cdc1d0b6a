// C ABI of libhelmholtz_amd: operator handle, shifted-Laplacian multigrid
// hierarchy, multigrid cycle and the batched flexible GMRES driver.
//
// Control flow follows solveLinearSystem (src/ShiftedLaplacianMultigridSolver.jl:33-102):
// zero-RHS shortcut, lazy hierarchy setup, outer operator = the true
// (unshifted) H, FGMRES preconditioned by one MG cycle, warning on max-iter.
// The hierarchy is the rediscretised one of the reference's commented
// geometric path (:51-62): shift folded into gamma, m and gamma restricted by
// full weighting, the operator rebuilt matrix-free on every coarse mesh.
#include <algorithm>
#include <chrono>
#include <map>
#include <atomic>
#include <mutex>
#include <thread>
#include <tuple>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/helmholtz_amd.h"
#include "hh_abl.h"
#include "hh_kernels.h"
#include "hh_unet.h"

using namespace hh;

namespace {

thread_local std::string g_err;

struct HHError : std::runtime_error {
  int code;
  HHError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hipck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HHError(HH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCK(x) hipck((x), #x)

template <typename T>
T* dalloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw HHError(HH_ERR_NOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
  }
  return (T*)p;
}
inline void dfree(void* p) {
  if (p) (void)hipFree(p);
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return HH_OK;
  } catch (const HHError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return HH_ERR_HIP;
  }
}

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

Grid make_grid(int dim, const int64_t nn[3], const double h[3], int bc) {
  Grid g;
  g.dim = dim;
  g.n1 = (int)nn[0];
  g.n2 = (int)nn[1];
  g.n3 = dim == 3 ? (int)nn[2] : 1;
  g.N = (int64_t)g.n1 * g.n2 * g.n3;
  g.ih1 = 1.0 / (h[0] * h[0]);
  g.ih2 = 1.0 / (h[1] * h[1]);
  g.ih3 = dim == 3 ? 1.0 / (h[2] * h[2]) : 0.0;
  g.bc = bc;
  return g;
}

// Absorbing-layer profile from per-axis 1-D ramps (src/GetHelmholtz.jl:97-220):
// 2-D  gamma = amp (p1 + p2 - p1 p2);  3-D  gamma = min(amp, amp (g1 + g2 + g3)).
__global__ void k_abl(int dim, int n1, int n2, int n3, const double* p1, const double* p2, const double* p3,
                      double amp, double* out) {
  const int64_t N = (int64_t)n1 * n2 * n3;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  const int i = (int)(p % n1);
  const int64_t q = p / n1;
  const int j = (int)(q % n2);
  const int k = (int)(q / n2);
  if (dim == 2) {
    const double a = p1[i], b = p2[j];
    out[p] = (a + b - a * b) * amp;
  } else {
    const double g = amp * (p1[i] + p2[j] + p3[k]);
    out[p] = g >= amp ? amp : g;
  }
}

// Sommerfeld boundary terms getSommerfeldBC (src/GetHelmholtz.jl:222-247):
// S = -i w (bc / h_d) sqrt(m) summed over the faces a node lies on, in the
// reference's face order (last axis first, its low face skipped with a Neumann
// top), each term formed as (w (bc / h_d)) sqrt(m).
__global__ void k_sommerfeld(int dim, int n1, int n2, int n3, double h1, double h2, double h3,
                             const double* __restrict__ m, double w, int neumann_top, double bc,
                             cx<double>* __restrict__ out) {
  const int64_t N = (int64_t)n1 * n2 * n3;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  const int i = (int)(p % n1);
  const int64_t q = p / n1;
  const int j = (int)(q % n2);
  const int k = (int)(q / n2);
  const double sm = sqrt(m[p]);
  const int ii[3] = {i, j, k}, nn[3] = {n1, n2, n3};
  const double hh[3] = {h1, h2, h3};
  double s = 0.0;
  for (int d = dim - 1; d >= 0; --d) {
    const double t = (w * (bc / hh[d])) * sm;
    const bool top = d == dim - 1;
    if (ii[d] == 0 && !(top && neumann_top)) s += t;
    if (ii[d] == nn[d] - 1) s += t;
  }
  out[p] = mk<double>(0.0, -s);
}

__global__ void k_sqrt_arr(const double* a, double* b, int n) {
  int i = blockIdx.x * 64 + threadIdx.x;
  if (i < n) b[i] = sqrt(a[i]);
}

}  // namespace

void hh_set_last_error(const char* msg) { g_err = msg ? msg : ""; }

namespace {
// per-level network features: k^2 h^2 = Re(w)^2 m h_x^2 and gamma / Re(w)
__global__ void k_level_feats(const double* __restrict__ m, const double* __restrict__ gam, int64_t N, double wre,
                              double hx2, float* __restrict__ kh2, float* __restrict__ gw) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  kh2[p] = (float)(wre * wre * m[p] * hx2);
  gw[p] = (float)(gam[p] / wre);
}
__global__ void k_fill_f(float* __restrict__ a, float v, int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p < n) a[p] = v;
}
__global__ void k_zero_cols(hh::cx<double>* __restrict__ x, int64_t N, const int* __restrict__ bad) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  if (bad[blockIdx.y]) x[(int64_t)blockIdx.y * N + p] = hh::mk<double>(0.0, 0.0);
}
}  // namespace

// ============================================================== operator
struct hh_op {
  int dev = 0;  // the GPU the operator's memory and stream live on
  int dim = 2;
  int64_t nn[3] = {1, 1, 1};
  double h[3] = {1, 1, 1};
  double wre = 0, wim = 0;
  int neumann_top = 0, sommerfeld = 0, bc = 2;
  // spread operator of GetHelmholtzOperatorHO (bc = 1 for its multigrid hierarchy)
  int ho = 0;
  double beta_lap = 1.0, beta_mass = 1.0;
  Grid g;
  double* m = nullptr;
  double* gam = nullptr;
  cx<double>* d64 = nullptr;  // diagonal of H, complex128
  cx<float>* d32 = nullptr;   // diagonal of H, complex64
  hipStream_t stream = nullptr;
  void* stage_in = nullptr;   // host-pointer staging
  void* stage_out = nullptr;
  size_t stage_bytes = 0;

  void ensure_stage(size_t bytes) {
    if (bytes <= stage_bytes) return;
    dfree(stage_in);
    dfree(stage_out);
    stage_in = stage_out = nullptr;
    stage_bytes = 0;
    stage_in = dalloc<char>(bytes);
    stage_out = dalloc<char>(bytes);
    stage_bytes = bytes;
  }
  ~hh_op() {
    dfree(m);
    dfree(gam);
    dfree(d64);
    dfree(d32);
    dfree(stage_in);
    dfree(stage_out);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// ============================================================== solver
namespace {

struct Level {
  Grid g;
  double h[3];
  double* m = nullptr;
  double* gam = nullptr;
  void* d = nullptr;      // diagonal of the shifted operator, preconditioner precision
  void* f = nullptr;      // damped-Jacobi factor relax (d + Laplacian diagonal)^-1, same precision
  void* xa = nullptr;     // ping-pong iterate buffers (cap * N)
  void* xb = nullptr;
  void* b = nullptr;      // right-hand side of this level (levels >= 1)
  void* r = nullptr;      // residual scratch
  void* cur = nullptr;    // buffer holding the current iterate
  // K-cycle (levels that receive a Krylov-accelerated correction): c1, A c1, scalars
  void* kc1 = nullptr;
  void* kv1 = nullptr;
  cx<double>* ksc = nullptr;
  // Jac-GMRES smoothing: Krylov basis of the level (steps + 1 blocks)
  void* gv = nullptr;
};

}  // namespace

struct hh_solver {
  hh_op* op = nullptr;
  hh_mg_params p{};
  std::vector<double> shift;
  // hierarchy
  bool built = false;
  std::vector<Level> lv;
  cx<double>* coarse_aug = nullptr;  // n x 2n augmented; right half = inverse
  cx<double>* coarse_d = nullptr;    // coarsest diagonal in complex128
  int64_t* piv = nullptr;
  cx<double>* fac = nullptr;
  cx<double>* pinv = nullptr;
  float* coarse_pack = nullptr;      // fp32 real-ified inverse for the MFMA apply (complex64 mode)
  int cap = 0;  // RHS capacity of the work buffers
  // Krylov storage
  cx<double>* V = nullptr;  // (inner+1) blocks
  void* Z = nullptr;        // inner blocks in preconditioner precision
  cx<double>* xk = nullptr; // iterate
  cx<double>* bk = nullptr; // right-hand sides
  cx<double>* wk = nullptr; // H z / residual
  void* b0 = nullptr;       // fine-level V-cycle input in c64 mode
  KrylovWork kw{};
  int* skip_copy = nullptr;
  cx<double>* bsc = nullptr;  // BiCGSTAB per-RHS scalars (rho, alpha, omega, beta)
  cx<double>* kpart = nullptr; // K-cycle partial inner products
  // Jac-GMRES smoothing: per-RHS complex128 state shared by the levels
  cx<double>* gm_part = nullptr;
  cx<double>* gm_hc = nullptr;
  cx<double>* gm_H = nullptr;
  double* gm_sc = nullptr;
  double* gm_beta = nullptr;
  cx<double>* gm_y = nullptr;
  int* conv_d = nullptr;
  double* n2tmp = nullptr;
  double* res_d = nullptr;
  int* h_done = nullptr;   // pinned host mirrors
  int* h_kact = nullptr;
  double* h_res = nullptr;
  double setup_time = 0, solve_time = 0;
  int64_t n_prec = 0;
  // learned coarse-grid correction
  hh_unet* net = nullptr;
  int learn_level = -1;
  bool use_learned = false;
  bool fell_back = false;
  bool learned_ran = false;  // the network was applied in the last solve
  // outcome of the learned/plain trial on this hierarchy and network:
  // -1 not run yet, 0 plain cycle kept, 1 network kept (later solves skip the trial)
  int learn_verdict = -1;
  // HH_LEARN_TRIAL: the two-restart trial decides; HH_LEARN_ALWAYS: the network
  // runs in every restart cycle (only non-finite results switch to the plain cycle)
  int learn_policy = HH_LEARN_TRIAL;
  float* feat_kh2 = nullptr;
  float* feat_gw = nullptr;
  int feat_level = -1;       // level the feature fields were built for
  float* nn_scale = nullptr;
  float h2_learn = 1.f;
  // this solver's own inference workspace: activations, input and output
  // staging (the network's weights are shared read-only with other solvers)
  hh::UNetWork nn_ws;
  hh::bf16* nn_x16 = nullptr;
  float* nn_y = nullptr;
  int64_t nn_pix = 0;
  // launch shapes picked by the autotuner (0 = heuristic), per level
  int rows_tuned[16] = {0};
  int tuned_for = 0;
  // hipGraph of one FGMRES restart cycle (replayed from the second cycle on)
  hipGraphExec_t gcycle = nullptr;
  int gkey_nrhs = 0;
  bool gkey_learned = false;
  bool gkey_tr = false;
  int gkey_zun = 0;
  // the outer operator is the spread operator's transpose (adjoint solve in progress)
  bool ho_tr = false;
  void drop_graphs() {
    if (gcycle) (void)hipGraphExecDestroy(gcycle);
    gcycle = nullptr;
  }
  // capture of level-l right-hand sides for training pairs
  int cap_level = -1;
  int64_t cap_max = 0, cap_count = 0;
  int cap_lo = 0, cap_hi = 0;
  bool cap_now = false;
  cx<float>* cap_buf = nullptr;
  // per-kernel timing (hh_solver_time_kernel); kt_active = RHS still iterating
  KTimer kt;
  int kt_active = 0;

  void free_work() {
    drop_graphs();
    tuned_for = 0;
    for (auto& L : lv) {
      dfree(L.xa); dfree(L.xb); dfree(L.b); dfree(L.r); dfree(L.kc1); dfree(L.kv1); dfree(L.ksc); dfree(L.gv);
      L.xa = L.xb = L.b = L.r = L.cur = L.kc1 = L.kv1 = L.gv = nullptr;
      L.ksc = nullptr;
    }
    dfree(kpart);
    kpart = nullptr;
    dfree(gm_part); dfree(gm_hc); dfree(gm_H); dfree(gm_sc); dfree(gm_beta); dfree(gm_y);
    gm_part = gm_hc = gm_H = gm_y = nullptr;
    gm_sc = gm_beta = nullptr;
    dfree(V); dfree(Z); dfree(xk); dfree(bk); dfree(wk); dfree(b0);
    V = nullptr; Z = nullptr; xk = bk = wk = nullptr; b0 = nullptr;
    dfree(kw.part); dfree(kw.partn); dfree(kw.h1); dfree(kw.h2); dfree(kw.nrm); dfree(kw.H); dfree(kw.cs);
    dfree(kw.sn); dfree(kw.gv); dfree(kw.y); dfree(kw.bnorm); dfree(kw.est); dfree(kw.done); dfree(kw.kact);
    dfree(kw.vscale); dfree(kw.fix); dfree(kw.wn2); dfree(kw.h1s); dfree(bsc);
    bsc = nullptr;
    kw = KrylovWork{};
    dfree(skip_copy); dfree(conv_d); dfree(n2tmp); dfree(res_d);
    skip_copy = nullptr; conv_d = nullptr; n2tmp = nullptr; res_d = nullptr;
    if (h_done) (void)hipHostFree(h_done);
    if (h_kact) (void)hipHostFree(h_kact);
    if (h_res) (void)hipHostFree(h_res);
    h_done = nullptr; h_kact = nullptr; h_res = nullptr;
    cap = 0;
  }
  void free_hier() {
    free_work();
    for (auto& L : lv) { dfree(L.m); dfree(L.gam); dfree(L.d); dfree(L.f); }
    lv.clear();
    dfree(coarse_aug); dfree(coarse_d); dfree(piv); dfree(fac); dfree(pinv); dfree(coarse_pack);
    coarse_aug = nullptr; coarse_d = nullptr; piv = nullptr; fac = nullptr; pinv = nullptr; coarse_pack = nullptr;
    built = false;
    learn_verdict = -1;
  }
  void free_learned() {
    dfree(feat_kh2); dfree(feat_gw); dfree(nn_scale); dfree(nn_x16); dfree(nn_y);
    feat_kh2 = nullptr; feat_gw = nullptr; nn_scale = nullptr; nn_x16 = nullptr; nn_y = nullptr;
    nn_pix = 0;
    feat_level = -1;
    nn_ws.release();
  }
  // input/output staging of the network for npix pixels (all images)
  void ensure_nn_stage(int64_t npix) {
    if (npix <= nn_pix) return;
    dfree(nn_x16); dfree(nn_y);
    nn_x16 = nullptr; nn_y = nullptr; nn_pix = 0;
    nn_x16 = dalloc<hh::bf16>((size_t)npix * hh::UNET_CIN);
    nn_y = dalloc<float>((size_t)npix * 2);
    nn_pix = npix;
  }
  ~hh_solver() {
    free_hier();
    free_learned();
    dfree(cap_buf);
  }
};

namespace {

size_t csz(int prec) { return prec == HH_C64 ? sizeof(cx<float>) : sizeof(cx<double>); }

constexpr int64_t kMaxCoarsest = 10000;  // nodes of the dense coarsest solve

void build_hierarchy(hh_solver* S) {
  hh_op* op = S->op;
  hipStream_t st = op->stream;
  const int L = S->p.levels;
  S->lv.resize(L);
  int64_t nn[3] = {op->nn[0], op->nn[1], op->dim == 3 ? op->nn[2] : 1};
  double h[3] = {op->h[0], op->h[1], op->h[2]};
  for (int l = 0; l < L; ++l) {
    Level& lv = S->lv[l];
    lv.g = make_grid(op->dim, nn, h, op->bc);
    memcpy(lv.h, h, sizeof(h));
    lv.m = dalloc<double>(lv.g.N);
    lv.gam = dalloc<double>(lv.g.N);
    if (l == 0) {
      HIPCK(hipMemcpyAsync(lv.m, op->m, sizeof(double) * lv.g.N, hipMemcpyDeviceToDevice, st));
      HIPCK(hipMemcpyAsync(lv.gam, op->gam, sizeof(double) * lv.g.N, hipMemcpyDeviceToDevice, st));
      launch_add_const(lv.gam, S->shift[0] * op->wre, lv.g.N, st);
    } else {
      Level& f = S->lv[l - 1];
      launch_restrict_real(f.g, lv.g, f.m, lv.m, st);
      launch_restrict_real(f.g, lv.g, f.gam, lv.gam, st);
      launch_add_const(lv.gam, (S->shift[l] - S->shift[l - 1]) * op->wre, lv.g.N, st);
    }
    if (S->p.prec == HH_C64) {
      lv.d = dalloc<cx<float>>(lv.g.N);
      launch_diag<float>(lv.g, lv.m, lv.gam, op->wre, op->wim, op->neumann_top, op->sommerfeld, lv.h,
                         (cx<float>*)lv.d, st);
      lv.f = dalloc<cx<float>>(lv.g.N);
      launch_jacobi_factor<float>(lv.g, lv.m, lv.gam, op->wre, op->wim, op->neumann_top, op->sommerfeld, lv.h,
                                  S->p.relax, (cx<float>*)lv.f, st);
    } else {
      lv.d = dalloc<cx<double>>(lv.g.N);
      launch_diag<double>(lv.g, lv.m, lv.gam, op->wre, op->wim, op->neumann_top, op->sommerfeld, lv.h,
                          (cx<double>*)lv.d, st);
      lv.f = dalloc<cx<double>>(lv.g.N);
      launch_jacobi_factor<double>(lv.g, lv.m, lv.gam, op->wre, op->wim, op->neumann_top, op->sommerfeld, lv.h,
                                   S->p.relax, (cx<double>*)lv.f, st);
    }
    if (l + 1 < L) {
      for (int d = 0; d < op->dim; ++d) {
        if ((nn[d] - 1) % 2 != 0)
          throw HHError(HH_ERR_ARG, "level " + std::to_string(l) + " has an odd cell count on axis " +
                                        std::to_string(d) + "; use fewer levels");
        nn[d] = (nn[d] - 1) / 2 + 1;
        h[d] *= 2.0;
        if (nn[d] < 3) throw HHError(HH_ERR_ARG, "too many levels for this grid");
      }
    }
  }
  // coarsest level: dense complex128 inverse of the shifted operator (Gauss-Jordan,
  // O(n^3) set-up and 32 n^2 bytes; 10^4 nodes admit the 129 x 65 coarsest grid of a
  // 2-level hierarchy on 257 x 129)
  Level& c = S->lv[L - 1];
  const int64_t n = c.g.N;
  if (n > kMaxCoarsest)
    throw HHError(HH_ERR_ARG, "coarsest grid has " + std::to_string(n) + " nodes (> " + std::to_string(kMaxCoarsest) +
                                  "); add levels");
  S->coarse_d = dalloc<cx<double>>(n);
  launch_diag<double>(c.g, c.m, c.gam, op->wre, op->wim, op->neumann_top, op->sommerfeld, c.h, S->coarse_d, st);
  S->coarse_aug = dalloc<cx<double>>((size_t)n * 2 * n);
  S->piv = dalloc<int64_t>(1);
  S->fac = dalloc<cx<double>>(n);
  S->pinv = dalloc<cx<double>>(1);
  launch_dense_inverse(c.g, S->coarse_d, S->coarse_aug, S->piv, S->fac, S->pinv, st);
  if (S->p.prec == HH_C64) {
    // complex64 preconditioner: the coarsest apply runs on MFMA from an fp32 copy
    S->coarse_pack = dalloc<float>((size_t)pack_inv_mfma_size(c.g.N));
    pack_inv_mfma(S->coarse_aug, c.g.N, S->coarse_pack, st);
  }
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(st));
  S->built = true;
}

// feature fields (k^2 h^2 and gamma / Re w) of grid level `level`, kept until
// another level asks for them (network inputs and training-pair features share
// the buffers, so they are rebuilt whenever the level differs)
void ensure_level_feats(hh_solver* S, int level) {
  if (S->feat_kh2 && S->feat_level == level) return;
  Level& L = S->lv[level];
  dfree(S->feat_kh2); dfree(S->feat_gw);
  S->feat_kh2 = S->feat_gw = nullptr;
  S->feat_level = -1;
  S->feat_kh2 = dalloc<float>(L.g.N);
  S->feat_gw = dalloc<float>(L.g.N);
  hipLaunchKernelGGL(k_level_feats, dim3((unsigned)((L.g.N + 255) / 256)), dim3(256), 0, S->op->stream, L.m, L.gam,
                     L.g.N, S->op->wre, L.h[0] * L.h[0], S->feat_kh2, S->feat_gw);
  HIPCK(hipGetLastError());
  S->feat_level = level;
}

// network features of the learned level (after the hierarchy exists)
void setup_learned(hh_solver* S) {
  if (!S->net) return;
  if (S->learn_level < 1 || S->learn_level >= (int)S->lv.size())
    throw HHError(HH_ERR_ARG, "learned level must be in 1..levels-1");
  Level& L = S->lv[S->learn_level];
  ensure_level_feats(S, S->learn_level);
  if (!S->nn_scale) S->nn_scale = dalloc<float>(65536);
  S->h2_learn = (float)(L.h[0] * L.h[0]);
}

// row-major A^-1 of the coarsest level (left half of the Gauss-Jordan workspace)
const cx<double>* coarse_inv(hh_solver* S) { return S->coarse_aug; }

// coarsest-level solve x = A^-1 r (MFMA in the complex64 mode)
template <typename T>
void coarse_apply(hh_solver* S, int64_t n, int nrhs, const cx<T>* r, cx<T>* x, const int* skip, hipStream_t st) {
  if (std::is_same<T, float>::value && S->coarse_pack)
    launch_dense_apply_mfma(S->coarse_pack, n, nrhs, (const cx<float>*)r, (cx<float>*)x, skip, st);
  else
    launch_dense_apply<T>(coarse_inv(S), n, nrhs, r, x, skip, st);
}

void ensure_work(hh_solver* S, int nrhs) {
  if (nrhs <= S->cap) return;
  S->free_work();
  const int pr = S->p.prec;
  const size_t es = csz(pr);
  for (size_t l = 0; l < S->lv.size(); ++l) {
    Level& L = S->lv[l];
    const size_t nb = (size_t)L.g.N * nrhs * es;
    L.xa = dalloc<char>(nb);
    L.xb = dalloc<char>(nb);
    if (l > 0) L.b = dalloc<char>(nb);
    L.r = dalloc<char>(nb);
    // K-cycle: levels 1 .. levels-2 receive the two-step Krylov correction
    if (S->p.cycle == 'K' && l >= 1 && l + 1 < S->lv.size()) {
      L.kc1 = dalloc<char>(nb);
      L.kv1 = dalloc<char>(nb);
      L.ksc = dalloc<cx<double>>((size_t)4 * nrhs);
    }
  }
  if (S->p.cycle == 'K') S->kpart = dalloc<cx<double>>((size_t)64 * 2 * nrhs);
  if (S->p.relax_type == HH_RELAX_JAC_GMRES) {
    for (size_t l = 0; l + 1 < S->lv.size(); ++l) {
      Level& L = S->lv[l];
      const int pl = S->p.per_level ? std::max(S->p.pre_l[l], S->p.post_l[l]) : std::max(S->p.pre, S->p.post);
      L.gv = dalloc<char>((size_t)L.g.N * nrhs * es * (pl + 1));
    }
    S->gm_part = dalloc<cx<double>>(gm_part_elems(nrhs));
    S->gm_hc = dalloc<cx<double>>((size_t)(GM_MAX_STEPS + 1) * nrhs);
    S->gm_H = dalloc<cx<double>>((size_t)nrhs * (GM_MAX_STEPS + 1) * GM_MAX_STEPS);
    S->gm_sc = dalloc<double>(nrhs);
    S->gm_beta = dalloc<double>(nrhs);
    S->gm_y = dalloc<cx<double>>((size_t)nrhs * GM_MAX_STEPS);
  }
  const int64_t N = S->lv[0].g.N;
  const int K = S->p.inner;
  const size_t blk = (size_t)N * nrhs;
  // BiCGSTAB keeps r, r^, p, v, s, t in the basis blocks and p^, s^ in Z
  const bool bicg = S->p.krylov == HH_BICGSTAB;
  // (a complex64 basis takes half the complex128 blocks)
  const bool vb32 = S->p.basis == HH_C64 && pr == HH_C64 && !S->op->ho;
  S->V = dalloc<cx<double>>(vb32 ? (blk * (K + 1) + 1) / 2 : blk * std::max(K + 1, bicg ? 6 : 0));
  S->Z = dalloc<char>(blk * std::max(K, bicg ? 2 : 0) * es);
  S->xk = dalloc<cx<double>>(blk);
  S->bk = dalloc<cx<double>>(blk);
  S->wk = dalloc<cx<double>>(blk);
  if (pr == HH_C64) S->b0 = dalloc<cx<float>>(blk);
  KrylovWork& w = S->kw;
  w.nrhs = nrhs;
  w.N = N;
  w.kmax = K;
  // node chunks per RHS for the Arnoldi/BiCGSTAB reductions: about 16K
  // blocks in all (64 per CU, so the last wave of blocks is a small tail),
  // chunks of at least 4096 nodes
  int nchunk = (int)std::max<int64_t>(1, std::min<int64_t>((16384 + nrhs - 1) / nrhs, (N + 4095) / 4096));
  // 3-D: chunk c is in-plane segment c mod nseg3 over run c / nseg3 of the
  // planes (runs of at least 8 planes) for the H z passes' z march
  w.nseg3 = 0;
  if (col3_ok(S->lv[0].g)) {
    const Grid& g0 = S->lv[0].g;
    const int nseg = col3_segments(g0);
    const int runs = std::max(1, std::min(nchunk / nseg, g0.n3 / 8));
    nchunk = nseg * runs;
    w.nseg3 = nseg;
  }
  w.nchunk = nchunk;
  w.part = dalloc<cx<double>>((size_t)nchunk * (K + 1) * nrhs);
  w.partn = dalloc<double>((size_t)nchunk * nrhs);
  w.h1 = dalloc<cx<double>>((size_t)(K + 1) * nrhs);
  w.h2 = dalloc<cx<double>>((size_t)(K + 1) * nrhs);
  w.nrm = dalloc<double>(nrhs);
  w.H = dalloc<cx<double>>((size_t)nrhs * (K + 1) * K);
  w.cs = dalloc<cx<double>>((size_t)nrhs * K);
  w.sn = dalloc<cx<double>>((size_t)nrhs * K);
  w.gv = dalloc<cx<double>>((size_t)nrhs * (K + 1));
  w.y = dalloc<cx<double>>((size_t)nrhs * K);
  w.bnorm = dalloc<double>(nrhs);
  w.est = dalloc<double>(nrhs);
  w.done = dalloc<int>(nrhs);
  w.kact = dalloc<int>(nrhs);
  w.lazy = pr == HH_C64;
  w.zunscaled = 0;
  w.stage = nullptr;
  w.vscale = dalloc<double>((size_t)(K + 1) * nrhs);
  w.fix = dalloc<int>(nrhs);
  w.wn2 = dalloc<double>(nrhs);
  w.h1s = dalloc<cx<double>>((size_t)(K + 1) * nrhs);
  S->bsc = dalloc<cx<double>>((size_t)4 * nrhs);
  S->skip_copy = dalloc<int>(nrhs);
  S->conv_d = dalloc<int>(nrhs);
  S->n2tmp = dalloc<double>(nrhs);
  S->res_d = dalloc<double>(nrhs);
  HIPCK(hipHostMalloc((void**)&S->h_done, sizeof(int) * nrhs, hipHostMallocDefault));
  HIPCK(hipHostMalloc((void**)&S->h_kact, sizeof(int) * nrhs, hipHostMallocDefault));
  HIPCK(hipHostMalloc((void**)&S->h_res, sizeof(double) * nrhs, hipHostMallocDefault));
  S->cap = nrhs;
}

// ---------------------------------------------------------------- autotuner
// Picks, per level, the run length each stencil thread marches (rows per
// thread) for the damped-Jacobi sweep at this nrhs, by timing the candidates
// with HIP events on the solver's stream.  Results are cached per process,
// keyed by (grid, nrhs, precision).
std::map<std::tuple<int, int, int, int, int>, int> g_tune_cache;
std::mutex g_tune_mu;

template <typename T>
int tune_level(hh_solver* S, Level& L, int nrhs) {
  hipStream_t st = S->op->stream;
  const auto key = std::make_tuple(L.g.n1, L.g.n2, L.g.n3, nrhs, (int)sizeof(T));
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) return it->second;
  }
  const int64_t n = L.g.N * (int64_t)nrhs;
  launch_zero<T>((cx<T>*)L.xa, n, st);
  launch_zero<T>((cx<T>*)L.r, n, st);
  hipEvent_t e0, e1;
  HIPCK(hipEventCreate(&e0));
  HIPCK(hipEventCreate(&e1));
  int best = 0;
  float best_ms = 1e30f;
  const int cand[] = {2, 4, 8, 16, 32};
  for (int rows : cand) {
    const int len = L.g.dim == 3 ? L.g.n3 : L.g.n2;
    if (rows > len) continue;
    auto go = [&] {
      launch_stencil<T, T, T, M_JACOBI>(L.g, nrhs, (const cx<T>*)L.xa, (const cx<T>*)L.r, (cx<T>*)L.xb,
                                        (const cx<T>*)L.d, nullptr, mk<double>(0, 0), 0.8, nullptr, st, rows,
                                        (const cx<T>*)L.f);
    };
    go();
    HIPCK(hipEventRecord(e0, st));
    for (int r = 0; r < 3; ++r) go();
    HIPCK(hipEventRecord(e1, st));
    HIPCK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) best_ms = ms, best = rows;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCK(hipGetLastError());
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = best;
  return best;
}

void autotune(hh_solver* S, int nrhs) {
  if (S->tuned_for == nrhs) return;
  for (size_t l = 0; l + 1 < S->lv.size() && l < 16; ++l) {
    Level& L = S->lv[l];
    // small launches are latency-bound: keep the heuristic there
    if (L.g.N * (int64_t)nrhs < (1 << 20)) {
      S->rows_tuned[l] = 0;
      continue;
    }
    S->rows_tuned[l] = S->p.prec == HH_C64 ? tune_level<float>(S, L, nrhs) : tune_level<double>(S, L, nrhs);
  }
  S->tuned_for = nrhs;
}

// ---------------------------------------------------------------- MG cycle
template <typename T>
struct Cycle {
  hh_solver* S;
  int nrhs;
  const int* skip;
  hipStream_t st;

  void* other(Level& L) { return L.cur == L.xa ? L.xb : L.xa; }
  // coarse-grid corrections per visit of level l (W: 2 on the top wlevels levels)
  int ncorr_at(int l) const {
    return (S->p.cycle == 'W' && (S->p.wlevels <= 0 || l < S->p.wlevels)) ? 2 : 1;
  }
  // fused 3-D up leg (last coarse correction + 2 post sweeps), complex64
  bool use_up3(const Level& L, int l) const {
    return std::is_same<T, float>::value && S->p.fused && S->p.cycle != 'K' &&
           S->p.relax_type == HH_RELAX_JAC && l + 1 < (int)S->lv.size() && post_at(l) == 2 && fused3_ok(L.g);
  }
  // fused 3-D down leg (2 pre sweeps from zero + the residual), complex64
  bool use_down3(const Level& L, int l) const {
    return std::is_same<T, float>::value && S->p.fused && S->p.cycle != 'K' &&
           S->p.relax_type == HH_RELAX_JAC && l + 1 < (int)S->lv.size() && pre_at(l) == 2 && fused3_ok(L.g);
  }
  // fused 3-D middle step of a W visit (prolongation + the next residual)
  bool use_mid3(const Level& L, int l) const {
    return std::is_same<T, float>::value && S->p.fused && S->p.cycle != 'K' &&
           S->p.relax_type == HH_RELAX_JAC && l + 1 < (int)S->lv.size() && fused3_ok(L.g);
  }
  // damped-Jacobi sweeps before / after the coarse correction on level l
  int pre_at(int l) const { return S->p.per_level ? S->p.pre_l[l] : S->p.pre; }
  int post_at(int l) const { return S->p.per_level ? S->p.post_l[l] : S->p.post; }
  // fused legs: 2-D mirror-ghost levels with a coarser level below and V/W(2,2)
  bool use_fused(const Level& L) const {
    const size_t l = &L - &S->lv[0];
    return S->p.fused && S->p.cycle != 'K' && S->p.relax_type == HH_RELAX_JAC && l + 1 < S->lv.size() &&
           pre_at((int)l) == 2 &&
           post_at((int)l) == 2 && fused_ok(L.g);
  }
  int rows_of(const Level& L) const {
    const size_t l = &L - &S->lv[0];
    return l < 16 ? S->rows_tuned[l] : 0;
  }

  // Jac-GMRES smoothing of L.cur in place (zero_x: L.cur is written from x = 0)
  void gm(Level& L, const cx<T>* b, int steps, bool zero_x) {
    GmArgs<T> a;
    a.g = L.g;
    a.nrhs = nrhs;
    a.b = b;
    a.x = (cx<T>*)L.cur;
    a.zero_x = zero_x;
    a.d = (const cx<T>*)L.d;
    a.f = (const cx<T>*)L.f;
    a.V = (cx<T>*)L.gv;
    a.u = (cx<T>*)L.r;
    a.part = S->gm_part;
    a.hc = S->gm_hc;
    a.H = S->gm_H;
    a.sc = S->gm_sc;
    a.beta = S->gm_beta;
    a.y = S->gm_y;
    a.skip = skip;
    gm_smooth<T>(a, steps, st);
  }
  // post-smoothing with the configured relaxation
  void post_smooth(Level& L, const cx<T>* b, int n) {
    if (S->p.relax_type == HH_RELAX_JAC_GMRES)
      gm(L, b, n, false);
    else
      sweeps(L, b, n);
  }
  void sweeps(Level& L, const cx<T>* b, int n) {
    for (int s = 0; s < n; ++s) {
      void* o = other(L);
      launch_stencil<T, T, T, M_JACOBI>(L.g, nrhs, (const cx<T>*)L.cur, b, (cx<T>*)o, (const cx<T>*)L.d, nullptr,
                                        mk<double>(0, 0), S->p.relax, skip, st, rows_of(L), (const cx<T>*)L.f);
      L.cur = o;
    }
  }

  // K-cycle coarse correction on level lc (right-hand side in C.b, result in
  // C.cur): two minimal-residual (GCR) steps preconditioned by the cycle on
  // level lc; C.b is overwritten by the second step's residual, C.r holds A c2.
  void krylov_correction(Level& C, int lc) {
    const int64_t N = C.g.N;
    const size_t bytes = sizeof(cx<T>) * N * nrhs;
    auto A = [&](const void* x, void* y) {
      launch_stencil<T, T, T, M_APPLY>(C.g, nrhs, (const cx<T>*)x, nullptr, (cx<T>*)y, (const cx<T>*)C.d, nullptr,
                                       mk<double>(0, 0), 0.0, skip, st, rows_of(C));
    };
    cx<T>* rb = (cx<T>*)C.b;
    cx<T>* c1 = (cx<T>*)C.kc1;
    cx<T>* v1 = (cx<T>*)C.kv1;
    C.cur = nullptr;
    run(lc, rb, true, nullptr);
    HIPCK(hipMemcpyAsync(c1, C.cur, bytes, hipMemcpyDeviceToDevice, st));
    A(c1, v1);
    kc_dots<T>(N, nrhs, v1, rb, v1, v1, 2, S->kpart, skip, st);
    kc_scalar(N, nrhs, S->kpart, 1, C.ksc, skip, st);
    kc_update<T>(N, nrhs, 0, C.ksc, rb, v1, nullptr, nullptr, skip, st);  // r2 = rc - a1 v1
    C.cur = nullptr;
    run(lc, rb, true, nullptr);
    cx<T>* c2 = (cx<T>*)C.cur;
    cx<T>* v2 = (cx<T>*)C.r;
    A(c2, v2);
    kc_dots<T>(N, nrhs, v1, v2, nullptr, nullptr, 1, S->kpart, skip, st);
    kc_scalar(N, nrhs, S->kpart, 2, C.ksc, skip, st);
    kc_update<T>(N, nrhs, 1, C.ksc, v2, v1, c2, c1, skip, st);  // v2 -= g v1, c2 -= g c1
    kc_dots<T>(N, nrhs, v2, rb, v2, v2, 2, S->kpart, skip, st);
    kc_scalar(N, nrhs, S->kpart, 3, C.ksc, skip, st);
    kc_update<T>(N, nrhs, 2, C.ksc, c2, c1, nullptr, nullptr, skip, st);  // e = a1 c1 + a2 c2
  }

  // Solve approximately A_l x = b; zero_init: start from x = 0 (then L.cur is
  // chosen by the caller through `first`), else continue from L.cur.
  void run(int l, const cx<T>* b, bool zero_init, void* first) {
    Level& L = S->lv[l];
    const int nl = (int)S->lv.size();
    if (std::is_same<T, float>::value) {
      if (S->cap_now && l == S->cap_level && zero_init) {
        S->cap_now = false;
        if (S->cap_count + nrhs <= S->cap_max) {
          HIPCK(hipMemcpyAsync(S->cap_buf + S->cap_count * L.g.N, b, sizeof(cx<float>) * L.g.N * nrhs,
                               hipMemcpyDeviceToDevice, st));
          S->cap_count += nrhs;
        }
      }
      if (S->use_learned && l == S->learn_level) {
        L.cur = first ? first : L.xa;
        // the network's weights are read-only here; activations and staging
        // are this solver's own, launches go to this solver's stream
        S->ensure_nn_stage(L.g.N * nrhs);
        unet_prep_input(nrhs, L.g.N, (const cx<float>*)b, S->feat_kh2, S->feat_gw, S->nn_x16, S->nn_scale, skip, st);
        unet_forward(S->net->net, S->nn_ws, st, nrhs, (int)L.g.n1, (int)L.g.n2, S->nn_x16, S->nn_y);
        unet_post_output(nrhs, L.g.N, S->nn_y, S->nn_scale, S->h2_learn, (cx<float>*)L.cur, skip, st);
        return;
      }
    }
    if (l == nl - 1) {
      if (!L.cur) L.cur = first ? first : L.xa;
      coarse_apply<T>(S, L.g.N, nrhs, b, (cx<T>*)L.cur, skip, st);
      return;
    }
    const int pre = pre_at(l), post = post_at(l);
    if (use_fused(L) && zero_init) {
      Level& C = S->lv[l + 1];
      const int ncorr = ncorr_at(l);
      L.cur = first ? first : L.xa;
      // algorithmic bytes per launch: per active RHS the fields each leg
      // reads and writes once (s bytes per value), plus the shared diagonal
      // (and, for the legs that smooth, the shared Jacobi factor)
      const double s = sizeof(cx<T>), Nf = (double)L.g.N, Nc = (double)C.g.N, act = S->kt_active;
      const bool td = S->kt.match(KT_DOWN, l), tm = S->kt.match(KT_MID, l), tu = S->kt.match(KT_UP, l);
      // with two corrections (W) the down leg stores no iterate: the middle
      // leg forms it again from b (L.cur only fixes the ping-pong parity)
      const bool w2 = ncorr == 2;
      if (td) S->kt.begin(st);
      launch_down_fused<T>(L.g, C.g, nrhs, b, (const cx<T>*)L.d, (const cx<T>*)L.f, S->p.relax,
                           w2 ? nullptr : (cx<T>*)L.cur, (cx<T>*)C.b, skip, st);
      if (td) S->kt.end(st, act * ((w2 ? 1 : 2) * Nf * s + Nc * s) + 2 * Nf * s);
      C.cur = nullptr;
      run(l + 1, (const cx<T>*)C.b, true, nullptr);
      if (w2) {
        void* o = other(L);
        if (tm) S->kt.begin(st);
        launch_mid_fused<T>(L.g, C.g, nrhs, (const cx<T>*)C.cur, b, (const cx<T>*)L.d, (const cx<T>*)L.f,
                            S->p.relax, (cx<T>*)o, (cx<T>*)C.b, skip, st);
        if (tm) S->kt.end(st, act * (2 * Nf * s + 2 * Nc * s) + 2 * Nf * s);
        L.cur = o;
        C.cur = nullptr;
        run(l + 1, (const cx<T>*)C.b, true, nullptr);
      }
      void* o = other(L);
      if (tu) S->kt.begin(st);
      launch_up_fused<T>(L.g, C.g, nrhs, (const cx<T>*)L.cur, (const cx<T>*)C.cur, b, (const cx<T>*)L.f,
                         S->p.relax, (cx<T>*)o, skip, st);
      if (tu) S->kt.end(st, act * (3 * Nf * s + Nc * s) + Nf * s);  // the up leg reads no diagonal
      L.cur = o;
      return;
    }
    // the fused 3-D down leg leaves the restriction of this level's residual
    // in the coarse level's right-hand side
    const bool down3 = zero_init && use_down3(L, l);
    // kernel timing of the fused 3-D legs: algorithmic bytes as for the 2-D legs
    const double s3 = sizeof(cx<float>), Nf3 = (double)L.g.N, act3 = S->kt_active;
    if (down3) {
      L.cur = first ? first : L.xa;
      Level& C = S->lv[l + 1];
      const bool td = S->kt.match(KT_DOWN, l);
      if (td) S->kt.begin(st);
      launch_down_fused3(L.g, C.g, nrhs, (const cx<float>*)b, (const cx<float>*)L.d, (const cx<float>*)L.f,
                         S->p.relax, (cx<float>*)L.cur, (cx<float>*)C.b, skip, st);
      // b in, x2 and b_c out; d and F shared
      if (td) S->kt.end(st, act3 * (2 * Nf3 * s3 + (double)C.g.N * s3) + 2 * Nf3 * s3);
    } else if (zero_init) {
      L.cur = first ? first : L.xa;
      if (pre > 0 && S->p.relax_type == HH_RELAX_JAC_GMRES) {
        gm(L, b, pre, true);
      } else if (pre > 0) {
        launch_jacobi0<T>(L.g, nrhs, b, (cx<T>*)L.cur, (const cx<T>*)L.f, skip, st);
        sweeps(L, b, pre - 1);
      } else {
        launch_zero<T>((cx<T>*)L.cur, L.g.N * nrhs, st);
      }
    } else if (S->p.relax_type == HH_RELAX_JAC_GMRES) {
      gm(L, b, pre, false);
    } else {
      sweeps(L, b, pre);
    }
    Level& C = S->lv[l + 1];
    if (S->p.cycle == 'K') {
      launch_stencil<T, T, T, M_RESID>(L.g, nrhs, (const cx<T>*)L.cur, b, (cx<T>*)L.r, (const cx<T>*)L.d, nullptr,
                                       mk<double>(0, 0), 0.0, skip, st, rows_of(L));
      launch_restrict<T>(L.g, C.g, nrhs, (const cx<T>*)L.r, (cx<T>*)C.b, skip, st);
      if (l + 2 == nl) {  // the coarse level is solved exactly: one step
        C.cur = nullptr;
        run(l + 1, (const cx<T>*)C.b, true, nullptr);
      } else {
        krylov_correction(C, l + 1);
      }
      launch_prolong_add<T>(L.g, C.g, nrhs, (const cx<T>*)C.cur, (cx<T>*)L.cur, skip, st);
      post_smooth(L, b, post);
      return;
    }
    // W: two coarse-grid corrections per visit, each from this level's
    // current residual (iterating the coarse cycle on the restricted residual
    // instead diverges with rediscretised coarse operators near kh ~ 2).
    const int ncorr = ncorr_at(l);
    const bool up3 = use_up3(L, l), mid3 = use_mid3(L, l);
    bool have_bc = down3;  // C.b holds the restriction of the current residual
    for (int c = 0; c < ncorr; ++c) {
      if (!have_bc) {
        launch_stencil<T, T, T, M_RESID>(L.g, nrhs, (const cx<T>*)L.cur, b, (cx<T>*)L.r, (const cx<T>*)L.d, nullptr,
                                         mk<double>(0, 0), 0.0, skip, st, rows_of(L));
        launch_restrict<T>(L.g, C.g, nrhs, (const cx<T>*)L.r, (cx<T>*)C.b, skip, st);
      }
      have_bc = false;
      C.cur = nullptr;
      run(l + 1, (const cx<T>*)C.b, true, nullptr);
      if (up3 && c + 1 == ncorr) break;  // the last correction is added by the fused up leg
      if (mid3 && c + 1 < ncorr) {       // correction, next residual and its restriction in one pass
        void* o = other(L);
        const bool tm = S->kt.match(KT_MID, l);
        if (tm) S->kt.begin(st);
        launch_mid_fused3(L.g, C.g, nrhs, (const cx<float>*)L.cur, (const cx<float>*)C.cur, (const cx<float>*)b,
                          (const cx<float>*)L.d, (cx<float>*)o, (cx<float>*)C.b, skip, st);
        // x and b in, x' out, the coarse correction in and b_c out; the diagonal shared
        if (tm) S->kt.end(st, act3 * (3 * Nf3 * s3 + 2 * (double)C.g.N * s3) + Nf3 * s3);
        L.cur = o;
        have_bc = true;
      } else {
        launch_prolong_add<T>(L.g, C.g, nrhs, (const cx<T>*)C.cur, (cx<T>*)L.cur, skip, st);
      }
    }
    if (up3) {
      void* o = other(L);
      const bool tu = S->kt.match(KT_UP, l);
      if (tu) S->kt.begin(st);
      launch_up_fused3(L.g, C.g, nrhs, (const cx<float>*)L.cur, (const cx<float>*)C.cur, (const cx<float>*)b,
                       (const cx<float>*)L.f, S->p.relax, (cx<float>*)o, skip, st);
      if (tu) S->kt.end(st, act3 * (3 * Nf3 * s3 + (double)C.g.N * s3) + Nf3 * s3);  // x, b in, x out; e_c; F
      L.cur = o;
      return;
    }
    post_smooth(L, b, post);
  }

  // z = M^-1 b on the fine level, result written to `out`
  void apply(const cx<T>* b, cx<T>* out) {
    Level& L0 = S->lv[0];
    const int nl = (int)S->lv.size();
    if (nl == 1) {
      L0.cur = out;
      coarse_apply<T>(S, L0.g.N, nrhs, b, out, skip, st);
      return;
    }
    // number of buffer writes on level 0: pre (incl. the zero-start sweep) +
    // post, or for the fused legs down + (corrections - 1) + up (3-D: each
    // fused leg one write, the middle steps one each)
    const int flips = use_fused(L0) ? ncorr_at(0) + 1
                                    : (use_down3(L0, 0) ? 1 : pre_at(0)) + (use_mid3(L0, 0) ? ncorr_at(0) - 1 : 0) +
                                          (use_up3(L0, 0) ? 1 : post_at(0));
    // choose the ping-pong pair (out, xa) so the last write lands in out;
    // the level's own xb is put back on every exit, exceptions included, so
    // free_work never sees the caller's block in its place
    struct Restore {
      Level& L;
      void* xb;
      ~Restore() {
        L.xb = xb;
        L.cur = nullptr;
      }
    } restore{L0, L0.xb};
    L0.xb = out;
    // (Jac-GMRES smoothing updates the iterate in place: start in out)
    void* first = (pre_at(0) == 0 || S->p.relax_type == HH_RELAX_JAC_GMRES) ? (void*)out
                                                                           : ((flips % 2 == 1) ? (void*)out : L0.xa);
    run(0, b, true, first);
    if (L0.cur != out)
      HIPCK(hipMemcpyAsync(out, L0.cur, sizeof(cx<T>) * L0.g.N * nrhs, hipMemcpyDeviceToDevice, st));
  }
};

// z = M^-1 (scale * v); scale (per RHS, lazy basis normalisation) may be null
template <typename T>
void precondition(hh_solver* S, int nrhs, const cx<double>* vj, void* zj, const int* skip, hipStream_t st,
                  const double* scale = nullptr, bool staged = false) {
  Cycle<T> c{S, nrhs, skip, st};
  const int64_t blk = S->lv[0].g.N * (int64_t)nrhs;
  if (std::is_same<T, float>::value) {
    if (staged) {
      // the complex64 input is already in b0 (written by the Arnoldi pass)
    } else if (scale) {
      KrylovWork w = S->kw;
      w.nrhs = nrhs;
      kr_cast_scaled<double>(w, vj, scale, (cx<float>*)S->b0, st);
    } else {
      launch_cast<double, float>(vj, (cx<float>*)S->b0, blk, st);
    }
    c.apply((const cx<T>*)S->b0, (cx<T>*)zj);
  } else {
    c.apply((const cx<T>*)vj, (cx<T>*)zj);
  }
}

// the outer operator: w = H z (z in the preconditioner precision, w complex128)
template <typename TI>
void h_apply(hh_solver* S, int nrhs, const cx<TI>* x, cx<double>* y, const int* skip, hipStream_t st) {
  hh_op* op = S->op;
  if (op->ho)
    launch_ho_stencil<TI, M_APPLY>(op->g, op->beta_lap, op->beta_mass, nrhs, x, nullptr, y, op->d64, nullptr,
                                   mk<double>(0, 0), skip, st, S->ho_tr ? 1 : 0);
  else
    launch_stencil<TI, double, double, M_APPLY>(op->g, nrhs, x, nullptr, y, op->d64, nullptr, mk<double>(0, 0), 0.0,
                                                skip, st);
}

void h_resid(hh_solver* S, int nrhs, const cx<double>* x, const cx<double>* b, cx<double>* r, hipStream_t st) {
  hh_op* op = S->op;
  if (op->ho)
    launch_ho_stencil<double, M_RESID>(op->g, op->beta_lap, op->beta_mass, nrhs, x, b, r, op->d64, nullptr,
                                       mk<double>(0, 0), nullptr, st, S->ho_tr ? 1 : 0);
  else
    launch_stencil<double, double, double, M_RESID>(op->g, nrhs, x, b, r, op->d64, nullptr, mk<double>(0, 0), 0.0,
                                                    nullptr, st);
}

void run_fgmres(hh_solver* S, int nrhs, hh_solve_stats* stats) {
  hipStream_t st = S->op->stream;
  KrylovWork w = S->kw;
  w.nrhs = nrhs;  // capacity may exceed nrhs; kernels index with nrhs-major strides
  const int64_t N = S->lv[0].g.N;
  const int64_t blk = N * (int64_t)nrhs;
  const int K = S->p.inner;
  const double tol = S->p.tol;
  const size_t zes = csz(S->p.prec);
  // complex64 basis (validated at create: FGMRES, complex64 cycle, 5/7-point)
  const bool vb32 = S->p.basis == HH_C64 && w.lazy && !S->op->ho;
  std::vector<int> iters(nrhs, 0), conv(nrhs, 0);
  // ||b|| per RHS
  kr_norm2(w, S->bk, S->n2tmp, nullptr, st);
  hipLaunchKernelGGL(k_sqrt_arr, dim3((nrhs + 63) / 64), dim3(64), 0, st, S->n2tmp, w.bnorm, nrhs);
  HIPCK(hipMemcpyAsync(S->h_res, w.bnorm, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
  HIPCK(hipStreamSynchronize(st));
  std::vector<double> bn(S->h_res, S->h_res + nrhs);
  for (int b = 0; b < nrhs; ++b) conv[b] = !(bn[b] > 0.0);
  HIPCK(hipMemsetAsync(S->xk, 0, sizeof(cx<double>) * blk, st));
  HIPCK(hipMemcpyAsync(S->wk, S->bk, sizeof(cx<double>) * blk, hipMemcpyDeviceToDevice, st));
  std::vector<double> res(nrhs, 1.0), prev(nrhs, 1.0);
  int64_t nprec = 0;
  const bool have_net = S->net != nullptr && S->p.prec == HH_C64;
  const bool always = S->learn_policy == HH_LEARN_ALWAYS;
  S->use_learned = have_net && (always || S->learn_verdict != 0);
  S->fell_back = have_net && !always && S->learn_verdict == 0;
  S->learned_ran = S->use_learned;
  int trial = S->use_learned && !always && S->learn_verdict < 0 ? 0 : -1;
  double rho_learned = 1.0;
  if (S->use_learned) setup_learned(S);
  for (int outer = 0; outer < S->p.max_outer; ++outer) {
    HIPCK(hipMemcpyAsync(S->conv_d, conv.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, st));
    // plain complex64 cycle: the Arnoldi pass stages the next input and Z
    // holds unscaled directions (the learned correction is not linear, so it
    // gets normalised inputs)
    // (a complex64 basis is that input itself: no staged copy)
    w.zunscaled = w.lazy && !S->use_learned;
    w.stage = w.zunscaled && !vb32 ? (cx<float>*)S->b0 : nullptr;
    if (vb32)
      kr_start_impl<float>(w, S->wk, (cx<float>*)S->V, tol, S->conv_d, S->n2tmp, st);
    else
      kr_start_impl<double>(w, S->wk, S->V, tol, S->conv_d, S->n2tmp, st);
    HIPCK(hipMemcpyAsync(S->h_done, w.done, sizeof(int) * nrhs, hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    int active = 0;
    for (int b = 0; b < nrhs; ++b) active += S->h_done[b] == 0;
    if (active == 0) {
      // every RHS already meets the tolerance at the cycle start
      break;
    }
    const int64_t nprec_cycle0 = nprec;
    // one Arnoldi step j: preconditioner cycle, H z and the Arnoldi passes;
    // the kernels skip RHS whose device-side done flag is set
    auto step = [&](int j) {
      const cx<double>* vj = vb32 ? nullptr : S->V + (int64_t)j * blk;
      const cx<float>* vj32 = (const cx<float>*)S->V + (int64_t)j * blk;
      void* zj = (char*)S->Z + (size_t)j * blk * zes;
      // H z: reads z (8 or 16 B per value), writes w (16 B), shared c128 diagonal
      const bool ta = S->kt.match(KT_APPLY, 0);
      const double Nf = (double)N;
      const double* vsj = w.lazy ? w.vscale + (int64_t)j * nrhs : nullptr;
      const bool staged = w.stage && j > 0;
      if (S->p.prec == HH_C64 && w.lazy && !S->op->ho && (!ta || vb32)) {
        // H z formed inside the Arnoldi passes (w is never stored)
        const HzArgs hz{S->op->g, (const cx<float>*)zj, S->op->d64};
        if (vb32) {
          // the stored complex64 V_j is the cycle's input (scaled first
          // for the learned correction, which is not linear)
          const cx<float>* in = vj32;
          if (!w.zunscaled) {
            kr_cast_scaled<float>(w, vj32, vsj, (cx<float>*)S->b0, st);
            in = (const cx<float>*)S->b0;
          }
          Cycle<float> c{S, nrhs, w.done, st};
          c.apply(in, (cx<float>*)zj);
          kr_arnoldi_impl<float>(w, (const cx<float>*)S->V, S->wk, j, tol, S->skip_copy, &S->kt, active, st,
                                 &hz);
        } else {
          precondition<float>(S, nrhs, vj, zj, w.done, st, vsj, staged);
          kr_arnoldi_impl<double>(w, S->V, S->wk, j, tol, S->skip_copy, &S->kt, active, st, &hz);
        }
        return;
      }
      if (S->p.prec == HH_C64) {
        precondition<float>(S, nrhs, vj, zj, w.done, st, vsj, staged);
        if (ta) S->kt.begin(st);
        h_apply<float>(S, nrhs, (const cx<float>*)zj, S->wk, w.done, st);
        if (ta) S->kt.end(st, active * Nf * 24.0 + Nf * 16.0);
        if (w.zunscaled) kr_scale_rhs(w, S->wk, vsj, st);
      } else {
        precondition<double>(S, nrhs, vj, zj, w.done, st);
        if (ta) S->kt.begin(st);
        h_apply<double>(S, nrhs, (const cx<double>*)zj, S->wk, w.done, st);
        if (ta) S->kt.end(st, active * Nf * 32.0 + Nf * 16.0);
      }
      kr_arnoldi_impl<double>(w, S->V, S->wk, j, tol, S->skip_copy, &S->kt, active, st);
    };
    // end of a restart cycle: x += Z y, then the true residual b - H x in
    // complex128 and its norm per RHS
    auto finish = [&] {
      if (S->p.prec == HH_C64)
        kr_update_x<float>(w, (const cx<float>*)S->Z, S->xk, st);
      else
        kr_update_x<double>(w, (const cx<double>*)S->Z, S->xk, st);
      h_resid(S, nrhs, S->xk, S->bk, S->wk, st);
      kr_true_res(w, S->wk, S->res_d, S->n2tmp, st);
    };
    // hipGraph mode: from the second restart cycle on, the whole cycle (K
    // steps, x update, true residual) is one captured graph replayed with a
    // single host synchronisation per restart; RHS that converge inside the
    // cycle are skipped by the kernels' done flags.  The first cycle runs
    // eagerly (it allocates the network workspace and may capture training
    // data), as do timed solves and solves that capture training pairs.
    const bool graph_cycle = S->p.graph && outer >= 1 && !S->cap_buf && S->kt.kernel == KT_NONE;
    if (graph_cycle) {
      S->cap_now = false;
      if (S->gkey_nrhs != nrhs || S->gkey_learned != S->use_learned || S->gkey_tr != S->ho_tr ||
          S->gkey_zun != w.zunscaled) {
        S->drop_graphs();
        S->gkey_nrhs = nrhs;
        S->gkey_learned = S->use_learned;
        S->gkey_tr = S->ho_tr;
        S->gkey_zun = w.zunscaled;
      }
      if (!S->gcycle) {
        hipGraph_t g = nullptr;
        HIPCK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int j = 0; j < K; ++j) step(j);
        finish();
        HIPCK(hipStreamEndCapture(st, &g));
        const hipError_t ie = hipGraphInstantiate(&S->gcycle, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCK(ie);
      }
      HIPCK(hipGraphLaunch(S->gcycle, st));
    } else {
      for (int j = 0; j < K; ++j) {
        S->cap_now = S->cap_buf && nprec_cycle0 + j >= S->cap_lo && nprec_cycle0 + j < S->cap_hi && active == nrhs;
        S->kt_active = active;
        step(j);
        HIPCK(hipMemcpyAsync(S->h_done, w.done, sizeof(int) * nrhs, hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
        active = 0;
        for (int b = 0; b < nrhs; ++b) active += S->h_done[b] == 0;
        if (active == 0) break;
      }
      finish();
    }
    HIPCK(hipMemcpyAsync(S->h_kact, w.kact, sizeof(int) * nrhs, hipMemcpyDeviceToHost, st));
    HIPCK(hipMemcpyAsync(S->h_res, S->res_d, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    HIPCK(hipGetLastError());
    int nconv = 0;
    bool stagnated = false;
    std::vector<int> bad(nrhs, 0);
    // lock-step preconditioner applications of this cycle: the steps some RHS
    // still took
    int kmax_cycle = 0;
    for (int b = 0; b < nrhs; ++b)
      if (!conv[b]) kmax_cycle = std::max(kmax_cycle, S->h_kact[b]);
    nprec += kmax_cycle;
    for (int b = 0; b < nrhs; ++b) {
      iters[b] += conv[b] ? 0 : S->h_kact[b];
      res[b] = S->h_res[b];
      if (!conv[b]) {
        if (!(res[b] == res[b])) bad[b] = 1, stagnated = true;
        else if (res[b] > 0.9 * prev[b]) stagnated = true;
      }
    }
    // selection between the learned and the plain cycle, judged by the true
    // residual: restart cycle 0 runs the network, cycle 1 the plain cycle; the
    // network is kept only if its cycle reduced the residual clearly more
    // (geometric mean over the active RHS, 0.8x margin since a first cycle
    // from b reduces most)
    double lred = 0.0;
    int nact = 0;
    for (int b = 0; b < nrhs; ++b)
      if (!conv[b] && res[b] == res[b] && res[b] > 0.0 && prev[b] > 0.0) {
        lred += std::log(res[b] / prev[b]);
        ++nact;
      }
    const double rho = nact ? std::exp(lred / nact) : 0.0;
    bool bad_any = false;
    for (int b = 0; b < nrhs; ++b) bad_any |= bad[b] != 0;
    if (trial == 0 && !bad_any) {
      rho_learned = rho;
      S->use_learned = false;  // cycle 1: plain
      trial = 1;
      stagnated = false;
    } else if (trial == 1) {
      S->use_learned = rho_learned < 0.8 * rho;
      S->fell_back = !S->use_learned;
      S->learn_verdict = S->use_learned ? 1 : 0;
      if (S->p.verbose)
        fprintf(stderr, "[helmholtz_amd] restart reduction learned %.3f plain %.3f -> %s\n", rho_learned, rho,
                S->use_learned ? "learned" : "plain");
      trial = 2;
      stagnated = false;
    }
    if (S->use_learned && stagnated && (!always || bad_any)) {
      // the learned correction stopped helping: fall back to the plain cycle;
      // columns it drove to non-finite values restart from zero
      S->use_learned = false;
      S->fell_back = true;
      S->learn_verdict = 0;
      trial = 2;
      if (S->p.verbose) fprintf(stderr, "[helmholtz_amd] learned correction stagnated; plain cycle from restart %d\n", outer);
      HIPCK(hipMemcpyAsync(S->conv_d, bad.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_zero_cols, dim3((unsigned)((N + 255) / 256), nrhs), dim3(256), 0, st, S->xk, N, S->conv_d);
      h_resid(S, nrhs, S->xk, S->bk, S->wk, st);
      for (int b = 0; b < nrhs; ++b)
        if (bad[b]) res[b] = 1.0;
    }
    for (int b = 0; b < nrhs; ++b) {
      if (!conv[b] && res[b] <= tol) conv[b] = 1;
      if (!(res[b] == res[b])) conv[b] = 1;  // non-finite: stop iterating it
      nconv += conv[b];
      prev[b] = res[b];
    }
    if (S->p.verbose) fprintf(stderr, "[helmholtz_amd] restart %d: %d/%d RHS converged\n", outer, nconv, nrhs);
    if (nconv == nrhs) break;
  }
  S->n_prec += nprec;
  S->kt.flush(st);
  if (stats) {
    int mx = 0, nc = 0;
    for (int b = 0; b < nrhs; ++b) {
      int status;
      if (!(bn[b] > 0.0))
        status = HH_ZERO_RHS, res[b] = 0.0;
      else if (!(res[b] == res[b]))
        status = HH_NONFINITE;
      else if (res[b] <= tol)
        status = HH_CONVERGED;
      else
        status = HH_MAXITER;
      if (stats->status) stats->status[b] = status;
      if (stats->iters) stats->iters[b] = iters[b];
      if (stats->true_res) stats->true_res[b] = res[b];
      mx = std::max(mx, iters[b]);
      nc += (status == HH_CONVERGED || status == HH_ZERO_RHS);
    }
    stats->max_iters = mx;
    stats->n_converged = nc;
    stats->n_prec = nprec;
    stats->learned_used = S->learned_ran;
    stats->fell_back = S->fell_back;
  }
  bool warn = false;
  for (int b = 0; b < nrhs; ++b) warn |= (bn[b] > 0.0 && !(res[b] <= tol));
  if (warn) fprintf(stderr, "WARNING: MG solver reached maximum iterations without convergence\n");
}

// Batched right-preconditioned BiCGSTAB (KrylovMethods.jl's bicgstb as called
// by solveBiCGSTAB_MG, src/ShiftedLaplacianMultigridSolver.jl:91-93): two
// preconditioner applications per iteration, at most max_outer iterations,
// stop on the recursive residual ||r|| <= tol ||b||; then the true residual
// b - H x is formed in complex128 and RHS that do not meet the tolerance
// restart from it (r^ = r) while iterations remain.
void run_bicgstab(hh_solver* S, int nrhs, hh_solve_stats* stats) {
  hipStream_t st = S->op->stream;
  KrylovWork w = S->kw;
  w.nrhs = nrhs;
  w.lazy = 0;
  const int64_t N = S->lv[0].g.N;
  const int64_t blk = N * (int64_t)nrhs;
  const double tol = S->p.tol;
  const int maxit = S->p.max_outer;
  std::vector<int> iters(nrhs, 0), conv(nrhs, 0);
  kr_norm2(w, S->bk, S->n2tmp, nullptr, st);
  hipLaunchKernelGGL(k_sqrt_arr, dim3((nrhs + 63) / 64), dim3(64), 0, st, S->n2tmp, w.bnorm, nrhs);
  HIPCK(hipMemcpyAsync(S->h_res, w.bnorm, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
  HIPCK(hipStreamSynchronize(st));
  std::vector<double> bn(S->h_res, S->h_res + nrhs), res(nrhs, 1.0);
  for (int b = 0; b < nrhs; ++b) conv[b] = !(bn[b] > 0.0);
  HIPCK(hipMemsetAsync(S->xk, 0, sizeof(cx<double>) * blk, st));
  cx<double>* r = S->V;
  cx<double>* rh = S->V + blk;
  cx<double>* pv = S->V + 2 * blk;
  cx<double>* v = S->V + 3 * blk;
  cx<double>* sv = S->V + 4 * blk;
  cx<double>* t = S->V + 5 * blk;
  const size_t zes = csz(S->p.prec);
  void* ph = S->Z;
  void* shat = (char*)S->Z + (size_t)blk * zes;
  S->use_learned = false;
  S->learned_ran = false;
  S->fell_back = false;
  int64_t nprec = 0;
  int total = 0;
  auto prec_apply = [&](const cx<double>* in, void* z, cx<double>* out) {
    if (S->p.prec == HH_C64) {
      precondition<float>(S, nrhs, in, z, w.done, st);
      h_apply<float>(S, nrhs, (const cx<float>*)z, out, w.done, st);
    } else {
      precondition<double>(S, nrhs, in, z, w.done, st);
      h_apply<double>(S, nrhs, (const cx<double>*)z, out, w.done, st);
    }
  };
  for (int cycle = 0;; ++cycle) {
    // true residual in complex128
    h_resid(S, nrhs, S->xk, S->bk, r, st);
    kr_true_res(w, r, S->res_d, S->n2tmp, st);
    HIPCK(hipMemcpyAsync(S->h_res, S->res_d, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    int active = 0;
    for (int b = 0; b < nrhs; ++b) {
      res[b] = bn[b] > 0.0 ? S->h_res[b] : 0.0;
      if (!conv[b] && (res[b] <= tol || !(res[b] == res[b]))) conv[b] = 1;
      active += !conv[b];
    }
    if (S->p.verbose) fprintf(stderr, "[helmholtz_amd] BiCGSTAB cycle %d: %d/%d RHS converged\n", cycle, nrhs - active, nrhs);
    if (active == 0 || total >= maxit || cycle >= 20) break;
    HIPCK(hipMemcpyAsync(rh, r, sizeof(cx<double>) * blk, hipMemcpyDeviceToDevice, st));
    HIPCK(hipMemcpyAsync(S->conv_d, conv.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, st));
    bicg_start(w, S->conv_d, st);
    const int limit = maxit - total;
    int steps = 0;
    for (int it = 0; it < limit; ++it) {
      const int first = it == 0;
      bicg_dots(w, rh, r, nullptr, nullptr, 1, st);
      bicg_scalar(w, S->bsc, 0, first, tol, limit, st);
      bicg_p(w, S->bsc, r, pv, v, first, st);
      prec_apply(pv, ph, v);
      bicg_dots(w, rh, v, nullptr, nullptr, 1, st);
      bicg_scalar(w, S->bsc, 1, first, tol, limit, st);
      bicg_s(w, S->bsc, r, v, sv, st);
      prec_apply(sv, shat, t);
      bicg_dots(w, t, sv, t, t, 2, st);
      bicg_scalar(w, S->bsc, 2, first, tol, limit, st);
      if (S->p.prec == HH_C64)
        bicg_xr<float>(w, S->bsc, (const cx<float>*)ph, (const cx<float>*)shat, sv, t, S->xk, r, st);
      else
        bicg_xr<double>(w, S->bsc, (const cx<double>*)ph, (const cx<double>*)shat, sv, t, S->xk, r, st);
      bicg_scalar(w, S->bsc, 3, first, tol, limit, st);
      nprec += 2;
      ++steps;
      HIPCK(hipMemcpyAsync(S->h_done, w.done, sizeof(int) * nrhs, hipMemcpyDeviceToHost, st));
      HIPCK(hipStreamSynchronize(st));
      int left = 0;
      for (int b = 0; b < nrhs; ++b) left += S->h_done[b] == 0;
      if (left == 0) break;
    }
    HIPCK(hipMemcpyAsync(S->h_kact, w.kact, sizeof(int) * nrhs, hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    HIPCK(hipGetLastError());
    for (int b = 0; b < nrhs; ++b) iters[b] += conv[b] ? 0 : S->h_kact[b];
    total += steps;
  }
  S->n_prec += nprec;
  if (stats) {
    int mx = 0, nc = 0;
    for (int b = 0; b < nrhs; ++b) {
      int status;
      if (!(bn[b] > 0.0))
        status = HH_ZERO_RHS;
      else if (!(res[b] == res[b]))
        status = HH_NONFINITE;
      else if (res[b] <= tol)
        status = HH_CONVERGED;
      else
        status = HH_MAXITER;
      if (stats->status) stats->status[b] = status;
      if (stats->iters) stats->iters[b] = iters[b];
      if (stats->true_res) stats->true_res[b] = res[b];
      mx = std::max(mx, iters[b]);
      nc += (status == HH_CONVERGED || status == HH_ZERO_RHS);
    }
    stats->max_iters = mx;
    stats->n_converged = nc;
    stats->n_prec = nprec;
    stats->learned_used = 0;
    stats->fell_back = 0;
  }
  bool warn = false;
  for (int b = 0; b < nrhs; ++b) warn |= (bn[b] > 0.0 && !(res[b] <= tol));
  if (warn) fprintf(stderr, "WARNING: MG solver reached maximum iterations without convergence\n");
}

// getABL on the device: host ramps (hh_abl.h), combined per node by k_abl
void abl_impl(int dim, const int64_t nnodes[3], int neumann_at_first_dim, const int64_t pad[3], double amp,
              const int32_t* faces, double* out, int loc) {
  if (!nnodes || !pad || !out) throw HHError(HH_ERR_ARG, "null argument");
  if (dim != 2 && dim != 3) throw HHError(HH_ERR_ARG, "dim must be 2 or 3");
  if (loc != HH_HOST && loc != HH_DEVICE) throw HHError(HH_ERR_ARG, "loc must be HH_HOST or HH_DEVICE");
  std::vector<std::vector<double>> prof;
  try {
    prof = abl_profiles(dim, nnodes, neumann_at_first_dim, pad, faces);
  } catch (const std::invalid_argument& e) {
    throw HHError(HH_ERR_ARG, e.what());
  }
  const int64_t N = (int64_t)prof[0].size() * (int64_t)prof[1].size() * (int64_t)prof[2].size();
  double* dp[3] = {nullptr, nullptr, nullptr};
  double* dout = nullptr;
  hipError_t e = hipSuccess;
  try {
    for (int d = 0; d < 3; ++d) dp[d] = dalloc<double>(prof[d].size());
    dout = loc == HH_DEVICE ? out : dalloc<double>(N);
  } catch (...) {
    for (int d = 0; d < 3; ++d) dfree(dp[d]);
    throw;
  }
  for (int d = 0; d < 3 && e == hipSuccess; ++d)
    e = hipMemcpy(dp[d], prof[d].data(), sizeof(double) * prof[d].size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_abl, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, nullptr, dim, (int)prof[0].size(),
                       (int)prof[1].size(), (int)prof[2].size(), dp[0], dp[1], dp[2], amp, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess && loc == HH_HOST) e = hipMemcpy(out, dout, sizeof(double) * N, hipMemcpyDeviceToHost);
  for (int d = 0; d < 3; ++d) dfree(dp[d]);
  if (loc == HH_HOST) dfree(dout);
  hipck(e, "hh_abl");
}

}  // namespace

// ============================================================== C ABI
extern "C" {

const char* hh_last_error(void) { return g_err.c_str(); }
int hh_version(void) { return HH_API_VERSION; }

int hh_op_create(int dim, const int64_t nnodes[3], const double h[3], const double* m, const double* gamma,
                 double omega_re, double omega_im, int neumann_top, int sommerfeld, int bc_order, hh_op** out) {
  return guarded([&] {
    if (!out || !nnodes || !h || !m || !gamma) throw HHError(HH_ERR_ARG, "null argument");
    if (dim != 2 && dim != 3) throw HHError(HH_ERR_ARG, "dim must be 2 or 3");
    if (bc_order != 1 && bc_order != 2) throw HHError(HH_ERR_ARG, "bc_order must be 1 or 2");
    if (!(omega_re != 0.0)) throw HHError(HH_ERR_ARG, "Re(omega) must be nonzero");
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) {
      if (nnodes[d] < 3 || nnodes[d] > (1 << 24)) throw HHError(HH_ERR_ARG, "each axis needs 3..2^24 nodes");
      if (!(h[d] > 0.0)) throw HHError(HH_ERR_ARG, "h must be positive");
      N *= nnodes[d];
    }
    if (N > ((int64_t)1 << 31)) throw HHError(HH_ERR_ARG, "grid too large");
    hh_op* op = new hh_op();
    try {
      op->dim = dim;
      for (int d = 0; d < 3; ++d) {
        op->nn[d] = d < dim ? nnodes[d] : 1;
        op->h[d] = d < dim ? h[d] : 1.0;
      }
      op->wre = omega_re;
      op->wim = omega_im;
      op->neumann_top = neumann_top;
      op->sommerfeld = sommerfeld;
      op->bc = bc_order;
      op->g = make_grid(dim, op->nn, op->h, bc_order);
      HIPCK(hipGetDevice(&op->dev));
      HIPCK(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking));
      op->m = dalloc<double>(N);
      op->gam = dalloc<double>(N);
      op->d64 = dalloc<cx<double>>(N);
      op->d32 = dalloc<cx<float>>(N);
      HIPCK(hipMemcpyAsync(op->m, m, sizeof(double) * N, hipMemcpyHostToDevice, op->stream));
      HIPCK(hipMemcpyAsync(op->gam, gamma, sizeof(double) * N, hipMemcpyHostToDevice, op->stream));
      launch_diag<double>(op->g, op->m, op->gam, omega_re, omega_im, neumann_top, sommerfeld, op->h, op->d64,
                          op->stream);
      launch_diag<float>(op->g, op->m, op->gam, omega_re, omega_im, neumann_top, sommerfeld, op->h, op->d32,
                         op->stream);
      HIPCK(hipGetLastError());
      HIPCK(hipStreamSynchronize(op->stream));
    } catch (...) {
      delete op;
      throw;
    }
    *out = op;
  });
}

int hh_op_create_ho(int dim, const int64_t nnodes[3], const double h[3], const double* m, const double* gamma,
                    double omega_re, double omega_im, int neumann_top, int sommerfeld, double beta_lap,
                    double beta_mass, hh_op** out) {
  return guarded([&] {
    if (!(beta_lap == beta_lap) || !(beta_mass == beta_mass)) throw HHError(HH_ERR_ARG, "beta must be finite");
    hh_op* op = nullptr;
    // the multigrid hierarchy of this operator uses the natural-Neumann (BC = 1)
    // 5/7-point stencil, which is the spread operator at beta = 1
    const int rc = hh_op_create(dim, nnodes, h, m, gamma, omega_re, omega_im, neumann_top, sommerfeld, 1, &op);
    if (rc != HH_OK) throw HHError(rc, g_err);
    op->ho = 1;
    op->beta_lap = beta_lap;
    op->beta_mass = dim == 2 ? beta_lap : beta_mass;
    *out = op;
  });
}

int hh_op_destroy(hh_op* op) {
  return guarded([&] { delete op; });
}

// Staging scratch of the allocation-free operator apply: a fixed device array,
// input half then output half.  Each GPU the library runs on holds its own
// copy until the process ends.  In host code `scratch` names a host-side
// stand-in, never a device address: kernels and copies take the per-device
// address from scratch_ptr().
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr size_t kScratchHalf = kScratchBytes / 2;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_dev[kMaxDevices] = {};

// device address of `scratch` on GPU dev, looked up once per device with that
// device current (every GPU holds the array at its own address)
void* scratch_ptr(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw HHError(HH_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    int cur = 0;
    HIPCK(hipGetDevice(&cur));
    if (cur != dev) HIPCK(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    HIPCK(e);
    if (!ptr) throw HHError(HH_ERR_HIP, "no device address for the scratch array");
    g_scratch_dev[dev] = ptr;
  }
  return g_scratch_dev[dev];
}

// Y = (H + i shift w^2 m) X on the operator's stream, X and Y device blocks
void op_apply_device(hh_op* op, double shift, const void* xin, void* yout, int64_t nrhs, int prec) {
  // i * shift * omega^2
  const double w2r = op->wre * op->wre - op->wim * op->wim, w2i = 2.0 * op->wre * op->wim;
  const cx<double> sfac = mk<double>(-shift * w2i, shift * w2r);
  const double* mm = shift != 0.0 ? op->m : nullptr;
  if (op->ho) {
    if (prec != HH_C128) throw HHError(HH_ERR_UNSUPPORTED, "the spread (HO) operator applies in complex128");
    launch_ho_stencil<double, M_APPLY>(op->g, op->beta_lap, op->beta_mass, (int)nrhs, (const cx<double>*)xin,
                                       nullptr, (cx<double>*)yout, op->d64, mm, sfac, nullptr, op->stream);
  } else if (prec == HH_C128)
    launch_stencil<double, double, double, M_APPLY>(op->g, (int)nrhs, (const cx<double>*)xin, nullptr,
                                                    (cx<double>*)yout, op->d64, mm, sfac, 0.0, nullptr, op->stream);
  else
    launch_stencil<float, float, float, M_APPLY>(op->g, (int)nrhs, (const cx<float>*)xin, nullptr,
                                                 (cx<float>*)yout, op->d32, mm, sfac, 0.0, nullptr, op->stream);
  HIPCK(hipGetLastError());
}
}  // namespace

int hh_op_apply(hh_op* op, double shift, const void* X, void* Y, int64_t nrhs, int prec, int loc) {
  return guarded([&] {
    if (!op || !X || !Y) throw HHError(HH_ERR_ARG, "null argument");
    if (nrhs < 1 || nrhs > 65535) throw HHError(HH_ERR_ARG, "nrhs must be in 1..65535");
    if (prec != HH_C128 && prec != HH_C64) throw HHError(HH_ERR_ARG, "prec must be HH_C128 or HH_C64");
    if (loc != HH_HOST && loc != HH_DEVICE) throw HHError(HH_ERR_ARG, "loc must be HH_HOST or HH_DEVICE");
    const size_t bytes = (size_t)op->g.N * nrhs * csz(prec);
    const void* xin = X;
    void* yout = Y;
    if (loc == HH_HOST) {
      op->ensure_stage(bytes);
      HIPCK(hipMemcpyAsync(op->stage_in, X, bytes, hipMemcpyHostToDevice, op->stream));
      xin = op->stage_in;
      yout = op->stage_out;
    }
    op_apply_device(op, shift, xin, yout, nrhs, prec);
    if (loc == HH_HOST) HIPCK(hipMemcpyAsync(Y, yout, bytes, hipMemcpyDeviceToHost, op->stream));
    HIPCK(hipStreamSynchronize(op->stream));
  });
}

int hh_op_apply_noalloc(hh_op* op, double shift, const void* X, void* Y, int64_t nrhs, int prec) {
  // hh_op_apply's host path with the staging in the fixed scratch array: no
  // hipMalloc / hipFree here or in anything it calls.  The stencil writes
  // every output value, so nothing in the scratch needs clearing first.
  return guarded([&] {
    if (!op || !X || !Y) throw HHError(HH_ERR_ARG, "null argument");
    if (nrhs < 1 || nrhs > 65535) throw HHError(HH_ERR_ARG, "nrhs must be in 1..65535");
    if (prec != HH_C128 && prec != HH_C64) throw HHError(HH_ERR_ARG, "prec must be HH_C128 or HH_C64");
    const size_t bytes = (size_t)op->g.N * nrhs * csz(prec);
    if (bytes > kScratchHalf)
      throw HHError(HH_ERR_ARG, "N * nrhs * element size = " + std::to_string(bytes) +
                                    " bytes exceeds the allocation-free apply's limit of " +
                                    std::to_string(kScratchHalf) + " bytes per block");
    char* base = (char*)scratch_ptr(op->dev);
    void* xin = base;
    void* yout = base + kScratchHalf;
    HIPCK(hipMemcpyAsync(xin, X, bytes, hipMemcpyHostToDevice, op->stream));
    op_apply_device(op, shift, xin, yout, nrhs, prec);
    HIPCK(hipMemcpyAsync(Y, yout, bytes, hipMemcpyDeviceToHost, op->stream));
    HIPCK(hipStreamSynchronize(op->stream));
  });
}

int hh_solver_create(hh_op* op, const hh_mg_params* p, const double* shift_per_level, hh_solver** out) {
  return guarded([&] {
    if (!op || !p || !out || !shift_per_level) throw HHError(HH_ERR_ARG, "null argument");
    if (p->levels < 1 || p->levels > 16) throw HHError(HH_ERR_ARG, "levels must be in 1..16");
    if (p->inner < 1 || p->inner > 32) throw HHError(HH_ERR_ARG, "inner must be in 1..32");
    if (p->pre < 0 || p->post < 0 || p->pre > 64 || p->post > 64) throw HHError(HH_ERR_ARG, "bad sweep counts");
    if (p->per_level)
      for (int l = 0; l < p->levels; ++l)
        if (p->pre_l[l] < 0 || p->post_l[l] < 0 || p->pre_l[l] > 64 || p->post_l[l] > 64)
          throw HHError(HH_ERR_ARG, "bad per-level sweep counts");
    if (p->relax_type != HH_RELAX_JAC && p->relax_type != HH_RELAX_JAC_GMRES)
      throw HHError(HH_ERR_ARG, "relax_type must be HH_RELAX_JAC or HH_RELAX_JAC_GMRES");
    if (p->relax_type == HH_RELAX_JAC_GMRES) {
      bool ok = p->pre <= GM_MAX_STEPS && p->post <= GM_MAX_STEPS;
      if (p->per_level)
        for (int l = 0; l < p->levels; ++l) ok = ok && p->pre_l[l] <= GM_MAX_STEPS && p->post_l[l] <= GM_MAX_STEPS;
      if (!ok) throw HHError(HH_ERR_ARG, "Jac-GMRES smoothing takes at most 8 steps per call");
    }
    if (p->cycle != 'V' && p->cycle != 'W' && p->cycle != 'K') throw HHError(HH_ERR_ARG, "cycle must be 'V', 'W' or 'K'");
    if (p->krylov != HH_FGMRES && p->krylov != HH_BICGSTAB) throw HHError(HH_ERR_ARG, "unknown krylov method");
    if (p->prec != HH_C128 && p->prec != HH_C64) throw HHError(HH_ERR_ARG, "prec must be HH_C128 or HH_C64");
    if (p->basis != HH_C128 && p->basis != HH_C64) throw HHError(HH_ERR_ARG, "basis must be HH_C128 or HH_C64");
    if (p->basis == HH_C64 && (p->krylov != HH_FGMRES || p->prec != HH_C64 || op->ho))
      throw HHError(HH_ERR_ARG, "basis = HH_C64 needs krylov = HH_FGMRES, prec = HH_C64 and the 5/7-point operator");
    if (p->learned != 0) throw HHError(HH_ERR_UNSUPPORTED, "learned correction not available in this entry point");
    if (p->max_outer < 1 || !(p->tol > 0.0)) throw HHError(HH_ERR_ARG, "max_outer >= 1 and tol > 0 required");
    hh_solver* s = new hh_solver();
    s->op = op;
    s->p = *p;
    s->shift.assign(shift_per_level, shift_per_level + p->levels);
    *out = s;
  });
}

int hh_solver_copy(const hh_solver* s, hh_solver** out) {
  return guarded([&] {
    if (!s || !out) throw HHError(HH_ERR_ARG, "null argument");
    hh_solver* c = new hh_solver();
    c->op = s->op;
    c->p = s->p;
    c->shift = s->shift;
    *out = c;
  });
}

int hh_solver_clear(hh_solver* s) {
  return guarded([&] {
    if (!s) throw HHError(HH_ERR_ARG, "null argument");
    s->free_hier();
  });
}

int hh_solver_destroy(hh_solver* s) {
  return guarded([&] { delete s; });
}

int hh_solver_levels(const hh_solver* s, int* levels, int64_t nnodes_of_level[][3], int max_levels) {
  return guarded([&] {
    if (!s || !levels) throw HHError(HH_ERR_ARG, "null argument");
    *levels = s->built ? (int)s->lv.size() : 0;
    if (nnodes_of_level)
      for (int l = 0; l < *levels && l < max_levels; ++l) {
        nnodes_of_level[l][0] = s->lv[l].g.n1;
        nnodes_of_level[l][1] = s->lv[l].g.n2;
        nnodes_of_level[l][2] = s->lv[l].g.n3;
      }
  });
}

int hh_solve(hh_solver* s, const void* B, void* X, int64_t nrhs, int do_transpose, int loc, hh_solve_stats* stats) {
  return guarded([&] {
    if (!s || !B || !X) throw HHError(HH_ERR_ARG, "null argument");
    if (nrhs < 1 || nrhs > 65535) throw HHError(HH_ERR_ARG, "nrhs must be in 1..65535");
    if (do_transpose != 0 && do_transpose != 1) throw HHError(HH_ERR_ARG, "do_transpose must be 0 or 1");
    if (loc != HH_HOST && loc != HH_DEVICE) throw HHError(HH_ERR_ARG, "loc must be HH_HOST or HH_DEVICE");
    hh_op* op = s->op;
    hipStream_t st = op->stream;
    const int64_t N = op->g.N;
    const size_t bytes = sizeof(cx<double>) * N * nrhs;
    double t0 = now_s();
    if (!s->built) build_hierarchy(s);
    ensure_work(s, (int)nrhs);
    if (s->p.autotune) autotune(s, (int)nrhs);
    double t1 = now_s();
    s->setup_time += t1 - t0;
    if (stats) stats->setup_time = t1 - t0;
    HIPCK(hipMemcpyAsync(s->bk, B, bytes, loc == HH_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, st));
    const int nr = (int)nrhs;
    KrylovWork w = s->kw;
    w.nrhs = nr;
    s->ho_tr = false;
    std::vector<double> bn0;
    const double tol0 = s->p.tol;
    if (do_transpose) {
      // adjoint solve (doTranspose = 1, src/ShiftedLaplacianMultigridSolver.jl:68-80): H^H x = b
      // <=> H y = W^-1 conj(b), x = W conj(y).  ||W^-1|| <= 2^dim, so the
      // inner tolerance tol / 2^dim bounds ||H^H x - b|| <= tol ||b||.
      kr_norm2(w, s->bk, s->n2tmp, nullptr, st);
      bn0.resize(nr);
      HIPCK(hipMemcpyAsync(bn0.data(), s->n2tmp, sizeof(double) * nr, hipMemcpyDeviceToHost, st));
      adjoint_map(op->g, s->bk, nr, 0, st);
      if (op->bc == 2) s->p.tol = tol0 / (double)(1 << op->dim);
      // spread operator (its hierarchy has bc = 1, so W = I and the map is a
      // conjugation): H^H x = b <=> H^T y = conj(b), x = conj(y), and
      // ||H^T y - conj(b)|| = ||H^H x - b||.  The 5-point preconditioner is
      // complex symmetric, so it serves H^T as it serves H.
      s->ho_tr = op->ho != 0;
    }
    try {
      if (s->p.krylov == HH_BICGSTAB)
        run_bicgstab(s, nr, stats);
      else
        run_fgmres(s, nr, stats);
    } catch (...) {
      s->p.tol = tol0;
      s->ho_tr = false;
      throw;
    }
    s->p.tol = tol0;
    if (do_transpose) {
      // adjoint residual W conj(c - H y) and the solution x = W conj(y)
      h_resid(s, nr, s->xk, s->bk, s->wk, st);
      adjoint_map(op->g, s->wk, nr, 2, st);
      kr_norm2(w, s->wk, s->n2tmp, nullptr, st);
      std::vector<double> rn(nr);
      HIPCK(hipMemcpyAsync(rn.data(), s->n2tmp, sizeof(double) * nr, hipMemcpyDeviceToHost, st));
      adjoint_map(op->g, s->xk, nr, 1, st);
      s->ho_tr = false;
      HIPCK(hipStreamSynchronize(st));
      if (stats) {
        int nc = 0;
        for (int b = 0; b < nr; ++b) {
          const double res = bn0[b] > 0.0 ? sqrt(rn[b] / bn0[b]) : 0.0;
          int status = !(bn0[b] > 0.0) ? HH_ZERO_RHS
                       : !(res == res) ? HH_NONFINITE
                       : res <= tol0   ? HH_CONVERGED
                                       : HH_MAXITER;
          if (stats->true_res) stats->true_res[b] = res;
          if (stats->status) stats->status[b] = status;
          nc += (status == HH_CONVERGED || status == HH_ZERO_RHS);
        }
        stats->n_converged = nc;
      }
    }
    HIPCK(hipMemcpyAsync(X, s->xk, bytes, loc == HH_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, st));
    HIPCK(hipStreamSynchronize(st));
    double t2 = now_s();
    s->solve_time += t2 - t1;
    if (stats) stats->solve_time = t2 - t1;
  });
}

int hh_sommerfeld_bc(int dim, const int64_t nnodes[3], const double h[3], const double* m, double omega,
                     int neumann_top, int bc_order, void* out, int loc) {
  return guarded([&] {
    if (!nnodes || !h || !m || !out) throw HHError(HH_ERR_ARG, "null argument");
    if (dim != 2 && dim != 3) throw HHError(HH_ERR_ARG, "dim must be 2 or 3");
    if (bc_order != 1 && bc_order != 2) throw HHError(HH_ERR_ARG, "bc_order must be 1 or 2");
    if (loc != HH_HOST && loc != HH_DEVICE) throw HHError(HH_ERR_ARG, "loc must be HH_HOST or HH_DEVICE");
    int64_t nn[3] = {1, 1, 1};
    double hv[3] = {1.0, 1.0, 1.0};
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) {
      if (nnodes[d] < 3 || nnodes[d] > (1 << 24)) throw HHError(HH_ERR_ARG, "each axis needs 3..2^24 nodes");
      if (!(h[d] > 0.0)) throw HHError(HH_ERR_ARG, "h must be positive");
      nn[d] = nnodes[d];
      hv[d] = h[d];
      N *= nn[d];
    }
    if (N > ((int64_t)1 << 31)) throw HHError(HH_ERR_ARG, "grid too large");
    const double* dm = m;
    cx<double>* dout = (cx<double>*)out;
    if (loc == HH_HOST) {
      double* t = dalloc<double>(N);
      dm = t;
      try {
        dout = dalloc<cx<double>>(N);
      } catch (...) {
        dfree(t);
        throw;
      }
      hipError_t e = hipMemcpy(t, m, sizeof(double) * N, hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        dfree(t);
        dfree(dout);
        hipck(e, "hh_sommerfeld_bc");
      }
    }
    hipLaunchKernelGGL(k_sommerfeld, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, nullptr, dim, (int)nn[0],
                       (int)nn[1], (int)nn[2], hv[0], hv[1], hv[2], dm, omega, neumann_top, (double)bc_order, dout);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && loc == HH_HOST) e = hipMemcpy(out, dout, sizeof(cx<double>) * N, hipMemcpyDeviceToHost);
    if (loc == HH_HOST) {
      dfree((double*)dm);
      dfree(dout);
    }
    hipck(e, "hh_sommerfeld_bc");
  });
}

int hh_abl(int dim, const int64_t nnodes[3], int neumann_at_first_dim, const int64_t pad[3], double amp,
           double* out, int loc) {
  return guarded([&] { abl_impl(dim, nnodes, neumann_at_first_dim, pad, amp, nullptr, out, loc); });
}

int hh_abl_faces(int dim, const int64_t nnodes[3], int neumann_at_first_dim, const int64_t pad[3], double amp,
                 const int32_t faces[6], double* out, int loc) {
  return guarded([&] {
    if (!faces) throw HHError(HH_ERR_ARG, "null argument");
    if (dim != 2 && dim != 3) throw HHError(HH_ERR_ARG, "dim must be 2 or 3");
    for (int q = 0; q < 2 * dim; ++q)
      if (faces[q] != 0 && faces[q] != 1) throw HHError(HH_ERR_ARG, "faces entries must be 0 or 1");
    abl_impl(dim, nnodes, neumann_at_first_dim, pad, amp, faces, out, loc);
  });
}

int hh_device_alloc(int64_t bytes, void** ptr) {
  return guarded([&] {
    if (!ptr || bytes < 0) throw HHError(HH_ERR_ARG, "bad argument");
    *ptr = dalloc<char>((size_t)bytes);
  });
}

int hh_device_free(void* ptr) {
  return guarded([&] { dfree(ptr); });
}

int hh_memcpy(void* dst, const void* src, int64_t bytes, int kind) {
  return guarded([&] {
    if (!dst || !src || bytes < 0) throw HHError(HH_ERR_ARG, "bad argument");
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    if (kind < 0 || kind > 2) throw HHError(HH_ERR_ARG, "kind must be 0 (H2D), 1 (D2H) or 2 (D2D)");
    HIPCK(hipMemcpy(dst, src, (size_t)bytes, k));
  });
}

int hh_op_time_kernel(hh_op* op, int kernel, double shift, const void* X, const void* B, void* Y, int64_t nrhs,
                      int prec, int reps, double* avg_ms) {
  return guarded([&] {
    if (!op || !X || !Y || !avg_ms || reps < 1) throw HHError(HH_ERR_ARG, "bad argument");
    if (nrhs < 1 || nrhs > 65535) throw HHError(HH_ERR_ARG, "nrhs must be in 1..65535");
    if (kernel < 0 || kernel > 2) throw HHError(HH_ERR_ARG, "kernel must be 0 (apply), 1 (residual), 2 (Jacobi)");
    if (kernel > 0 && !B) throw HHError(HH_ERR_ARG, "residual/Jacobi need B");
    if (op->ho) throw HHError(HH_ERR_UNSUPPORTED, "kernel timing covers the 5/7-point operator");
    const double w2r = op->wre * op->wre - op->wim * op->wim, w2i = 2.0 * op->wre * op->wim;
    const cx<double> sfac = mk<double>(-shift * w2i, shift * w2r);
    const double* mm = shift != 0.0 ? op->m : nullptr;
    const int n = (int)nrhs;
    auto launch = [&]() {
      if (prec == HH_C128) {
        auto x = (const cx<double>*)X;
        auto b = (const cx<double>*)B;
        auto y = (cx<double>*)Y;
        if (kernel == 0) launch_stencil<double, double, double, M_APPLY>(op->g, n, x, b, y, op->d64, mm, sfac, 0.8, nullptr, op->stream);
        if (kernel == 1) launch_stencil<double, double, double, M_RESID>(op->g, n, x, b, y, op->d64, mm, sfac, 0.8, nullptr, op->stream);
        if (kernel == 2) launch_stencil<double, double, double, M_JACOBI>(op->g, n, x, b, y, op->d64, mm, sfac, 0.8, nullptr, op->stream);
      } else {
        auto x = (const cx<float>*)X;
        auto b = (const cx<float>*)B;
        auto y = (cx<float>*)Y;
        if (kernel == 0) launch_stencil<float, float, float, M_APPLY>(op->g, n, x, b, y, op->d32, mm, sfac, 0.8, nullptr, op->stream);
        if (kernel == 1) launch_stencil<float, float, float, M_RESID>(op->g, n, x, b, y, op->d32, mm, sfac, 0.8, nullptr, op->stream);
        if (kernel == 2) launch_stencil<float, float, float, M_JACOBI>(op->g, n, x, b, y, op->d32, mm, sfac, 0.8, nullptr, op->stream);
      }
    };
    launch();  // warm-up
    hipEvent_t e0, e1;
    HIPCK(hipEventCreate(&e0));
    HIPCK(hipEventCreate(&e1));
    HIPCK(hipEventRecord(e0, op->stream));
    for (int r = 0; r < reps; ++r) launch();
    HIPCK(hipEventRecord(e1, op->stream));
    HIPCK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCK(hipGetLastError());
    *avg_ms = (double)ms / reps;
  });
}

int hh_solver_attach_unet(hh_solver* s, hh_unet* net, int level) {
  return guarded([&] {
    if (!s) throw HHError(HH_ERR_ARG, "null argument");
    if (net && s->p.prec != HH_C64) throw HHError(HH_ERR_ARG, "the learned correction needs the complex64 preconditioner");
    if (net && (level < 1 || level >= s->p.levels)) throw HHError(HH_ERR_ARG, "level must be in 1..levels-1");
    if (net && s->op->dim != 2) throw HHError(HH_ERR_UNSUPPORTED, "the learned correction is 2-D only");
    if (net) {
      int64_t n1 = s->op->nn[0], n2 = s->op->nn[1];
      for (int l = 0; l < level; ++l) n1 = (n1 - 1) / 2 + 1, n2 = (n2 - 1) / 2 + 1;
      if (n1 < 9 || n2 < 9 || (n1 - 1) % 8 || (n2 - 1) % 8)
        throw HHError(HH_ERR_ARG, "the learned level's grid needs 8k + 1 nodes per axis");
    }
    s->free_learned();
    s->learn_verdict = -1;
    s->net = net;
    s->learn_level = net ? level : -1;
  });
}

int hh_solver_learned_policy(hh_solver* s, int policy) {
  return guarded([&] {
    if (!s) throw HHError(HH_ERR_ARG, "null argument");
    if (policy != HH_LEARN_TRIAL && policy != HH_LEARN_ALWAYS)
      throw HHError(HH_ERR_ARG, "policy must be HH_LEARN_TRIAL or HH_LEARN_ALWAYS");
    s->learn_policy = policy;
    s->learn_verdict = -1;
    s->drop_graphs();
  });
}

int hh_solver_learned_workspace(const hh_solver* s, const void** activations, const void** staging) {
  return guarded([&] {
    if (!s || !activations || !staging) throw HHError(HH_ERR_ARG, "null argument");
    *activations = s->nn_ws.bufs.empty() ? nullptr : s->nn_ws.bufs[0];
    *staging = s->nn_x16;
  });
}

int hh_solver_capture_begin(hh_solver* s, int level, int64_t max_samples, int iter_lo, int iter_hi) {
  return guarded([&] {
    if (!s) throw HHError(HH_ERR_ARG, "null argument");
    if (s->p.prec != HH_C64) throw HHError(HH_ERR_ARG, "capture needs the complex64 preconditioner");
    if (level < 1 || level >= s->p.levels || max_samples < 1) throw HHError(HH_ERR_ARG, "bad level or size");
    if (!s->built) build_hierarchy(s);
    dfree(s->cap_buf);
    s->cap_buf = nullptr;
    s->cap_buf = dalloc<cx<float>>((size_t)max_samples * s->lv[level].g.N);
    s->cap_level = level;
    s->cap_max = max_samples;
    s->cap_count = 0;
    s->cap_lo = iter_lo;
    s->cap_hi = iter_hi;
  });
}

int hh_solver_capture_count(const hh_solver* s, int64_t* n) {
  return guarded([&] {
    if (!s || !n) throw HHError(HH_ERR_ARG, "null argument");
    *n = s->cap_count;
  });
}

int hh_dataset_add_from_solver(hh_dataset* d, hh_solver* s, int sweeps) {
  // Targets e = A_l^-1 r for the captured level-l residuals.  A multigrid
  // cycle started at a coarse level diverges there (kh ~ 2-4), so each block
  // is solved to 1e-6 by FGMRES(10) on the level's shifted operator, with a
  // W(2,2) preconditioner whose hierarchy carries a further 0.5 shift.
  return guarded([&] {
    if (!d || !s || sweeps < 1) throw HHError(HH_ERR_ARG, "bad argument");
    if (!s->cap_buf || s->cap_count == 0) throw HHError(HH_ERR_ARG, "solver has no captured samples");
    Level& L = s->lv[s->cap_level];
    if (L.g.n1 != d->nx || L.g.n2 != d->ny || L.g.dim != 2)
      throw HHError(HH_ERR_ARG, "dataset grid does not match the captured level");
    hipStream_t st = s->op->stream;
    HIPCK(hipStreamSynchronize(st));
    const int64_t N = L.g.N;
    std::vector<double> hm(N), hg(N);
    HIPCK(hipMemcpy(hm.data(), L.m, sizeof(double) * N, hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(hg.data(), L.gam, sizeof(double) * N, hipMemcpyDeviceToHost));
    const int64_t nn[3] = {L.g.n1, L.g.n2, 1};
    hh_op* op = nullptr;
    int rc = hh_op_create(2, nn, L.h, hm.data(), hg.data(), s->op->wre, s->op->wim, s->op->neumann_top,
                          s->op->sommerfeld, s->op->bc, &op);
    if (rc != HH_OK) throw HHError(rc, g_err);
    hh_solver* ts = nullptr;
    hh_mg_params tp{};
    tp.levels = (int)s->lv.size() - s->cap_level;
    tp.pre = 2;
    tp.post = 2;
    tp.relax = 0.8;
    tp.cycle = 'W';
    tp.max_outer = std::max(4, sweeps);
    tp.tol = 1e-6;
    tp.inner = 10;
    tp.krylov = HH_FGMRES;
    tp.prec = HH_C128;
    tp.fused = 1;
    std::vector<double> shifts(tp.levels, 0.5);
    cx<double>* rb = nullptr;
    cx<double>* xb = nullptr;
    try {
      rc = hh_solver_create(op, &tp, shifts.data(), &ts);
      if (rc != HH_OK) throw HHError(rc, g_err);
      const int chunk = 256;
      rb = dalloc<cx<double>>((size_t)N * chunk);
      xb = dalloc<cx<double>>((size_t)N * chunk);
      // features of the captured level (network inputs)
      ensure_level_feats(s, s->cap_level);
      for (int64_t k0 = 0; k0 < s->cap_count && d->n < d->cap;) {
        const int m = (int)std::min<int64_t>({(int64_t)chunk, s->cap_count - k0, d->cap - d->n});
        launch_cast<float, double>(s->cap_buf + k0 * N, rb, N * m, st);
        HIPCK(hipStreamSynchronize(st));
        hh_solve_stats stt{};
        rc = hh_solve(ts, rb, xb, m, 0, HH_DEVICE, &stt);
        if (rc != HH_OK) throw HHError(rc, g_err);
        HIPCK(hipMemcpyAsync(d->r + d->n * N, s->cap_buf + k0 * N, sizeof(cx<float>) * N * m,
                             hipMemcpyDeviceToDevice, st));
        launch_cast<double, float>(xb, d->e + d->n * N, N * m, st);
        for (int k = 0; k < m; ++k) {
          HIPCK(hipMemcpyAsync(d->kh2 + (d->n + k) * N, s->feat_kh2, sizeof(float) * N, hipMemcpyDeviceToDevice, st));
          HIPCK(hipMemcpyAsync(d->gw + (d->n + k) * N, s->feat_gw, sizeof(float) * N, hipMemcpyDeviceToDevice, st));
        }
        hipLaunchKernelGGL(k_fill_f, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, d->h2 + d->n,
                           (float)(L.h[0] * L.h[0]), (int64_t)m);
        HIPCK(hipStreamSynchronize(st));
        d->n += m;
        k0 += m;
      }
    } catch (...) {
      dfree(rb);
      dfree(xb);
      if (ts) hh_solver_destroy(ts);
      hh_op_destroy(op);
      throw;
    }
    dfree(rb);
    dfree(xb);
    hh_solver_destroy(ts);
    hh_op_destroy(op);
    if (s->net == nullptr) s->free_learned();
    HIPCK(hipGetLastError());
  });
}

int hh_stream_solve(int nprob, const hh_problem* probs, int nworkers, hh_problem_result* results) {
  // Uneven stream: independent 2-D/3-D problems of different sizes and RHS
  // counts.  Workers take problems largest-first (N * nrhs) from a shared
  // counter; each problem gets its own operator, solver and HIP stream, so
  // launches of different problems overlap on the device.  Results per
  // problem are bit-identical to solving it alone.
  return guarded([&] {
    if (nprob < 0 || (nprob > 0 && (!probs || !results))) throw HHError(HH_ERR_ARG, "bad argument");
    if (nworkers < 1 || nworkers > 16) throw HHError(HH_ERR_ARG, "nworkers must be in 1..16");
    std::vector<int> order(nprob);
    for (int i = 0; i < nprob; ++i) order[i] = i;
    auto work = [&](int i) {
      const hh_problem& p = probs[i];
      int64_t N = 1;
      for (int d = 0; d < p.dim; ++d) N *= p.nnodes[d];
      return (double)N * (double)std::max<int64_t>(1, p.nrhs);
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    std::atomic<int> next{0};
    std::mutex err_mu;
    std::string first_err;
    int first_code = HH_OK;
    auto worker = [&]() {
      for (;;) {
        const int k = next.fetch_add(1);
        if (k >= nprob) return;
        const int i = order[k];
        const hh_problem& p = probs[i];
        hh_problem_result& r = results[i];
        r.code = HH_OK;
        const double t0 = now_s();
        hh_op* op = nullptr;
        hh_solver* sv = nullptr;
        int rc = hh_op_create(p.dim, p.nnodes, p.h, p.m, p.gamma, p.omega_re, p.omega_im, p.neumann_top,
                              p.sommerfeld, p.bc_order, &op);
        if (rc == HH_OK) rc = hh_solver_create(op, &p.params, p.shift_per_level, &sv);
        if (rc == HH_OK && p.net) rc = hh_solver_attach_unet(sv, p.net, p.learn_level);
        if (rc == HH_OK) {
          hh_solve_stats st{};
          st.iters = p.iters;
          st.status = p.status;
          st.true_res = p.true_res;
          rc = hh_solve(sv, p.B, p.X, p.nrhs, 0, p.loc, &st);
          r.max_iters = st.max_iters;
          r.n_converged = st.n_converged;
          r.fell_back = st.fell_back;
        }
        if (rc != HH_OK) {
          r.code = rc;
          std::lock_guard<std::mutex> lk(err_mu);
          if (first_code == HH_OK) first_code = rc, first_err = g_err;
        }
        if (sv) hh_solver_destroy(sv);
        if (op) hh_op_destroy(op);
        r.seconds = now_s() - t0;
      }
    };
    std::vector<std::thread> th;
    for (int w = 0; w < std::min(nworkers, std::max(1, nprob)); ++w) th.emplace_back(worker);
    for (auto& t : th) t.join();
    if (first_code != HH_OK) throw HHError(first_code, "stream: " + first_err);
  });
}

int hh_solver_precondition(hh_solver* s, const void* V, void* Z, int64_t nrhs, int loc) {
  // One application of the multigrid preconditioner, Z = M^-1 V: V is
  // complex128, Z has the preconditioner's precision.  Test/inspection hook.
  return guarded([&] {
    if (!s || !V || !Z) throw HHError(HH_ERR_ARG, "null argument");
    if (nrhs < 1 || nrhs > 65535) throw HHError(HH_ERR_ARG, "nrhs must be in 1..65535");
    if (loc != HH_HOST && loc != HH_DEVICE) throw HHError(HH_ERR_ARG, "bad loc");
    if (!s->built) build_hierarchy(s);
    ensure_work(s, (int)nrhs);
    hipStream_t st = s->op->stream;
    const int64_t N = s->op->g.N;
    const size_t zb = csz(s->p.prec) * N * nrhs;
    cx<double>* v = s->V;  // first Krylov block as staging
    HIPCK(hipMemcpyAsync(v, V, sizeof(cx<double>) * N * nrhs,
                         loc == HH_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, st));
    if (s->p.prec == HH_C64)
      precondition<float>(s, (int)nrhs, v, s->Z, nullptr, st);
    else
      precondition<double>(s, (int)nrhs, v, s->Z, nullptr, st);
    HIPCK(hipMemcpyAsync(Z, s->Z, zb, loc == HH_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, st));
    HIPCK(hipStreamSynchronize(st));
    HIPCK(hipGetLastError());
  });
}

int hh_solver_time_kernel(hh_solver* s, int kernel, int level) {
  return guarded([&] {
    if (!s) throw HHError(HH_ERR_ARG, "null argument");
    if (kernel < HH_KT_NONE || kernel > HH_KT_UPDATE) throw HHError(HH_ERR_ARG, "unknown kernel id");
    if (kernel >= HH_KT_APPLY && level != 0) throw HHError(HH_ERR_ARG, "H z and the Arnoldi passes run on level 0");
    if (kernel == HH_KT_APPLY && s->p.basis == HH_C64)
      throw HHError(HH_ERR_UNSUPPORTED, "with the complex64 basis H z is formed inside the Arnoldi passes only");
    if (kernel >= HH_KT_DOWN && kernel <= HH_KT_UP && (level < 0 || level >= s->p.levels - 1))
      throw HHError(HH_ERR_ARG, "the fused legs run on levels 0..levels-2");
    s->kt.reset();
    s->kt.kernel = kernel;
    s->kt.level = level;
  });
}

int hh_solver_kernel_time(hh_solver* s, double* total_ms, int64_t* launches, double* alg_bytes) {
  return guarded([&] {
    if (!s || !total_ms || !launches || !alg_bytes) throw HHError(HH_ERR_ARG, "null argument");
    s->kt.flush(s->op->stream);
    *total_ms = s->kt.ms;
    *launches = s->kt.launches;
    *alg_bytes = s->kt.bytes;
  });
}

}  // extern "C"
