// Host-side launchers for the device kernels (internal to libhelmholtz_amd).
#pragma once
#include "hh_common.h"

#include <vector>

namespace hh {

enum { M_APPLY = 0, M_RESID = 1, M_JACOBI = 2 };

// y = A x | y = b - A x | y = x + wj D^-1 (b - A x);  A = Lap + diag(d + sfac m)
// rows_override > 0 fixes the run length each thread marches (autotuner), else a heuristic
template <typename TI, typename TO, typename TD, int MODE>
void launch_stencil(const Grid& g, int nrhs, const cx<TI>* x, const cx<TO>* b, cx<TO>* y, const cx<TD>* d,
                    const double* m, cx<double> sfac, double wj, const int* skip, hipStream_t s,
                    int rows_override = 0, const cx<TD>* fj = nullptr);
template <typename T>
void launch_zero(cx<T>* a, int64_t n, hipStream_t s);
template <typename T>
void launch_jacobi0(const Grid& g, int nrhs, const cx<T>* b, cx<T>* x, const cx<T>* f, const int* skip,
                    hipStream_t s);
template <typename T>
void launch_restrict(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* rf, cx<T>* bc, const int* skip,
                     hipStream_t s);
void launch_restrict_real(const Grid& gf, const Grid& gc, const double* f, double* c, hipStream_t s);
template <typename T>
void launch_prolong_add(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* xc, cx<T>* xf, const int* skip,
                        hipStream_t s);
template <typename A, typename B>
void launch_cast(const cx<A>* a, cx<B>* b, int64_t n, hipStream_t s);
template <typename T>
void launch_diag(const Grid& g, const double* m, const double* gam, double wre, double wim, int neumann_top,
                 int sommerfeld, const double h[3], cx<T>* d, hipStream_t s);
// f = relax * (d + Laplacian diagonal)^-1 (the damped-Jacobi factor), computed in double
template <typename T>
void launch_jacobi_factor(const Grid& g, const double* m, const double* gam, double wre, double wim, int neumann_top,
                          int sommerfeld, const double h[3], double relax, cx<T>* f, hipStream_t s);
void launch_add_const(double* a, double c, int64_t n, hipStream_t s);
// leaves A^-1 row-major in M[0, n*n) (and column-major in M[n*n, 2*n*n))
void launch_dense_inverse(const Grid& g, const cx<double>* d, cx<double>* M, int64_t* piv, cx<double>* fac,
                          cx<double>* pinv, hipStream_t s);
template <typename T>
void launch_dense_apply(const cx<double>* Ainv, int64_t n, int nrhs, const cx<T>* r, cx<T>* x, const int* skip,
                        hipStream_t s);
// complex64 coarsest apply on MFMA: P = the inverse packed by pack_inv_mfma (fp32, pack_inv_mfma_size(n) floats)
void pack_inv_mfma(const cx<double>* Ainv, int64_t n, float* P, hipStream_t s);
int64_t pack_inv_mfma_size(int64_t n);
void launch_dense_apply_mfma(const float* P, int64_t n, int nrhs, const cx<float>* r, cx<float>* x, const int* skip,
                             hipStream_t s);

// ---- spread (GetHelmholtzOperatorHO) operator (hh_ho.hip): M_APPLY y = H x (+ sfac m x), M_RESID y = b - H x ----
// tr = 1 applies the transpose H^T = Lap + D M (Lap and M are symmetric) in place of H = Lap + M D.
template <typename TI, int MODE>
void launch_ho_stencil(const Grid& g, double beta_lap, double beta_mass, int nrhs, const cx<TI>* x,
                       const cx<double>* b, cx<double>* y, const cx<double>* d, const double* m, cx<double> sfac,
                       const int* skip, hipStream_t s, int tr = 0);

// ---- K-cycle coarse correction (hh_kcycle.hip); part: 64*2*nrhs, sc: 4*nrhs ----
template <typename T>
void kc_dots(int64_t N, int nrhs, const cx<T>* a0, const cx<T>* b0, const cx<T>* a1, const cx<T>* b1, int nd,
             cx<double>* part, const int* skip, hipStream_t s);
void kc_scalar(int64_t N, int nrhs, const cx<double>* part, int stage, cx<double>* sc, const int* skip,
               hipStream_t s);
template <typename T>
void kc_update(int64_t N, int nrhs, int mode, const cx<double>* sc, cx<T>* y, const cx<T>* x, cx<T>* y2,
               const cx<T>* x2, const int* skip, hipStream_t s);

// ---- fused 2-D multigrid legs (hh_fused.hip): BC = 2, 2 pre/2 post sweeps ----
bool fused_ok(const Grid& gf);
// ---- fused 3-D legs (hh_fused3.hip), complex64, f = relax D^-1 ----
bool fused3_ok(const Grid& gf);
// down: x2 = two Jacobi sweeps from zero (into xo), b_c = R (b - A x2) (into bco)
void launch_down_fused3(const Grid& gf, const Grid& gc, int nrhs, const cx<float>* b, const cx<float>* d,
                        const cx<float>* f, double relax, cx<float>* xo, cx<float>* bco, const int* skip,
                        hipStream_t s);
// middle step of a W visit: x' = x + P e_c (into xo), b_c = R (b - A x') (into bco)
void launch_mid_fused3(const Grid& gf, const Grid& gc, int nrhs, const cx<float>* x, const cx<float>* ec,
                       const cx<float>* b, const cx<float>* d, cx<float>* xo, cx<float>* bco, const int* skip,
                       hipStream_t s);
// up: x' = x + P e_c, then two Jacobi sweeps (into xo)
void launch_up_fused3(const Grid& gf, const Grid& gc, int nrhs, const cx<float>* x, const cx<float>* ec,
                      const cx<float>* b, const cx<float>* f, double relax, cx<float>* xo, const int* skip,
                      hipStream_t s);
// f = relax D^-1 (the level's Jacobi factor) and relax itself;
// xo = null: the smoothed iterate is not stored (the middle leg recomputes it)
template <typename T>
void launch_down_fused(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* b, const cx<T>* d, const cx<T>* f,
                       double relax, cx<T>* xo, cx<T>* bco, const int* skip, hipStream_t s);
// the middle leg of a W visit: recomputes the down leg's x2 from b
template <typename T>
void launch_mid_fused(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* ec, const cx<T>* b, const cx<T>* d,
                      const cx<T>* f, double relax, cx<T>* xo, cx<T>* bco, const int* skip, hipStream_t s);
// the up leg's two sweeps need only f and relax (no diagonal)
template <typename T>
void launch_up_fused(const Grid& gf, const Grid& gc, int nrhs, const cx<T>* x, const cx<T>* ec, const cx<T>* b,
                     const cx<T>* f, double relax, cx<T>* xo, const int* skip, hipStream_t s);

// ---- per-kernel timing inside a solve ----
// Brackets every launch of one chosen kernel (at one grid level) with HIP
// events on the launching stream and accumulates the elapsed time, the launch
// count and the launches' algorithmic bytes.  Off while a graph is captured.
enum KtKernel { KT_NONE = 0, KT_DOWN = 1, KT_MID = 2, KT_UP = 3, KT_APPLY = 4, KT_DOTS = 5, KT_UPDATE = 6 };
struct KTimer {
  int kernel = KT_NONE, level = 0;
  bool enabled = true;
  std::vector<hipEvent_t> ev;
  std::vector<double> pend;  // algorithmic bytes of the recorded launches
  size_t used = 0;
  double ms = 0.0, bytes = 0.0;
  int64_t launches = 0;
  bool match(int k, int l) const { return enabled && kernel != KT_NONE && k == kernel && l == level; }
  void flush(hipStream_t s) {
    if (!used) return;
    (void)hipStreamSynchronize(s);
    for (size_t i = 0; i + 1 < used; i += 2) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, ev[i], ev[i + 1]) == hipSuccess) ms += t;
      bytes += pend[i / 2];
      ++launches;
    }
    used = 0;
    pend.clear();
  }
  void begin(hipStream_t s) {
    if (ev.empty()) {
      ev.resize(512);
      for (auto& e : ev) (void)hipEventCreate(&e);
    }
    if (used + 2 > ev.size()) flush(s);
    (void)hipEventRecord(ev[used], s);
  }
  void end(hipStream_t s, double b) {
    (void)hipEventRecord(ev[used + 1], s);
    pend.push_back(b);
    used += 2;
  }
  void reset() {
    used = 0;
    pend.clear();
    ms = bytes = 0.0;
    launches = 0;
  }
  ~KTimer() {
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
};

// ---- batched FGMRES (hh_krylov.hip) ----
struct KrylovWork {
  int nrhs;
  int64_t N;
  int kmax;              // restart length
  int nchunk;            // node chunks per RHS for the reductions
  int nseg3;             // 3-D H z passes: in-plane segments (nchunk = nseg3 x runs of planes); 0: linear chunks
  cx<double>* part;      // [nchunk][kmax+1][nrhs] partial dots
  double* partn;         // [nchunk][nrhs] partial squared norms
  cx<double>* h1;        // [kmax+1][nrhs]
  cx<double>* h2;        // [kmax+1][nrhs]
  double* nrm;           // [nrhs] scratch norm
  cx<double>* H;         // [nrhs][kmax+1][kmax] Hessenberg -> R
  cx<double>* cs;        // [nrhs][kmax]
  cx<double>* sn;        // [nrhs][kmax]
  cx<double>* gv;        // [nrhs][kmax+1]
  cx<double>* y;         // [nrhs][kmax]
  double* bnorm;         // [nrhs]
  double* est;           // [nrhs] residual estimate / ||b||
  int* done;             // [nrhs] 1 = no more work in this cycle
  int* kact;             // [nrhs] Krylov steps used by this RHS in this cycle
  // lazy normalisation: basis vector i of RHS b is vscale[i][b] * (stored V_i)
  int lazy;
  double* vscale;        // [kmax+1][nrhs]
  int* fix;              // [nrhs] 1 = the second CGS correction is applied this step
  double* wn2;           // [nrhs] ||w||^2 before projection (lazy mode)
  cx<double>* h1s;       // [kmax+1][nrhs] vscale[i] * h1[i] (lazy mode: coefficients of the stored vectors)
  // staged preconditioner input (lazy mode, plain cycle): pass 2 also writes
  // the complex64 copy of the stored w' into stage, which the next step's
  // cycle takes as its input; Z_j then holds M^-1 (stored V_j) and the true
  // z_j = vscale[j] Z_j (scale applied to H z and to the x update)
  int zunscaled;
  cx<float>* stage;      // [nrhs][N] or null
};
// b0 = (complex64) vscale_j[b] * V_j  (preconditioner input in the lazy mode;
// TV: basis element type)
template <typename TV>
void kr_cast_scaled(const KrylovWork& w, const cx<TV>* Vj, const double* scale, cx<float>* b0, hipStream_t s);
// w *= vscale_j[b] per RHS (H Z_j -> H z_j when Z holds unscaled directions)
void kr_scale_rhs(const KrylovWork& w, cx<double>* wv, const double* scale, hipStream_t s);

// ||v_b||^2 per RHS into out[b] (RHS with skip[b] != 0 untouched)
void kr_norm2(const KrylovWork& w, const cx<double>* v, double* out, const int* skip, hipStream_t s);
// Cycle start: beta_b = ||r_b||, V0 = r / beta, g = (beta, 0, ...), done/kact
// flags (conv[b] != 0 marks an RHS already converged).  n2tmp: nrhs doubles.
template <typename TV>
void kr_start_impl(const KrylovWork& w, const cx<double>* r, cx<TV>* V0, double tol, const int* conv,
                   double* n2tmp, hipStream_t s);
// w = H z formed inside the lazy Arnoldi passes: the outer 5/7-point
// operator's grid and complex128 diagonal, and the complex64 z of this step
struct HzArgs {
  Grid g;
  const cx<float>* z;
  const cx<double>* d;
};
// 3-D grids whose H z passes march z over in-plane segments of 512 nodes
// (n1 <= 768), and the number of such segments per plane
bool col3_ok(const Grid& g);
int col3_segments(const Grid& g);
// One CGS2 Arnoldi step j: orthogonalise wv against V_0..V_j, write V_{j+1},
// fill column j of H, update Givens rotations and residual estimates.
// skip_copy: nrhs ints of scratch.  hz (lazy mode only, may be null): take
// w = H z from hz instead of wv.  TV = float (complex64 basis) requires hz.
template <typename TV>
void kr_arnoldi_impl(const KrylovWork& w, const cx<TV>* V, cx<double>* wv, int j, double tol, int* skip_copy,
                     KTimer* kt, int active, hipStream_t s, const HzArgs* hz = nullptr);
// y = R^-1 g per RHS, x += Z y (Z in TZ)
template <typename TZ>
void kr_update_x(const KrylovWork& w, const cx<TZ>* Z, cx<double>* x, hipStream_t s);
// res[b] = ||r_b|| / ||b_b||
void kr_true_res(const KrylovWork& w, const cx<double>* r, double* res, double* n2tmp, hipStream_t s);

// ---- batched BiCGSTAB (hh_krylov.hip); sc: 4 * nrhs per-RHS scalars ----
void bicg_start(const KrylovWork& w, const int* conv, hipStream_t s);
void bicg_dots(const KrylovWork& w, const cx<double>* a0, const cx<double>* b0, const cx<double>* a1,
               const cx<double>* b1, int nd, hipStream_t s);
void bicg_scalar(const KrylovWork& w, cx<double>* sc, int stage, int first, double tol, int limit, hipStream_t s);
void bicg_p(const KrylovWork& w, const cx<double>* sc, const cx<double>* r, cx<double>* p, const cx<double>* v,
            int first, hipStream_t s);
void bicg_s(const KrylovWork& w, const cx<double>* sc, const cx<double>* r, const cx<double>* v, cx<double>* sv,
            hipStream_t s);
template <typename TZ>
void bicg_xr(const KrylovWork& w, const cx<double>* sc, const cx<TZ>* ph, const cx<TZ>* shat, const cx<double>* sv,
             const cx<double>* t, cx<double>* x, cx<double>* r, hipStream_t s);
// adjoint-solve transforms: mode 0 v = W^-1 conj(v), 1 v = W conj(v), 2 v = W v
void adjoint_map(const Grid& g, cx<double>* v, int nrhs, int mode, hipStream_t s);

// ---- Jac-GMRES smoothing (hh_gmsmooth.hip) ----
constexpr int GM_MAX_STEPS = 8;  // GMRES steps per smoothing call
// One level's operands and workspace: b, x (updated in place; zero_x: x = 0
// on entry and is written), diagonal d and damped-Jacobi factor f, basis V
// ((GM_MAX_STEPS + 1) blocks of N x nrhs), u (one block), and the per-RHS
// complex128 state: part (gm_part_elems), hc ((GM_MAX_STEPS + 1) x nrhs),
// H (nrhs x (GM_MAX_STEPS + 1) x GM_MAX_STEPS), sc, beta (nrhs), y
// (nrhs x GM_MAX_STEPS).
template <typename T>
struct GmArgs {
  Grid g;
  int nrhs;
  const cx<T>* b;
  cx<T>* x;
  bool zero_x;
  const cx<T>* d;
  const cx<T>* f;
  cx<T>* V;
  cx<T>* u;
  cx<double>* part;
  cx<double>* hc;
  cx<double>* H;
  double* sc;
  double* beta;
  cx<double>* y;
  const int* skip;
};
size_t gm_part_elems(int nrhs);
template <typename T>
void gm_smooth(const GmArgs<T>& a, int steps, hipStream_t st);

}  // namespace hh
