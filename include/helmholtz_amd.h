/*
 * helmholtz_amd.h — C ABI of libhelmholtz_amd.so, the MI355X (gfx950) path for
 * Helmholtz.jl's discrete Helmholtz operator and its shifted-Laplacian
 * multigrid (SL-MG) preconditioned Krylov solve.
 *
 * Plain C: opaque handles, int status codes (0 = success, message via
 * hh_last_error()), no exceptions cross the boundary.  Complex vectors are
 * interleaved (re, im) doubles (HH_C128) or floats (HH_C64), column-major
 * N x nrhs with N = n1*n2[*n3] nodes and x fastest — exactly Julia's
 * Matrix{ComplexF64}, so a Ptr{ComplexF64} passes straight through ccall.
 * "Top" is index 0 of the last axis.
 *
 * Threading: a handle must not be used by two host threads at once (the
 * reference solver holds shared scratch too, src/ShiftedLaplacianMultigridSolver.jl:74-75).
 * Every call is synchronous: it returns after its device work has finished.
 */
#ifndef HELMHOLTZ_AMD_H
#define HELMHOLTZ_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HH_API_VERSION 1

/* status codes */
enum { HH_OK = 0, HH_ERR_ARG = 1, HH_ERR_HIP = 2, HH_ERR_NOMEM = 3, HH_ERR_UNSUPPORTED = 4 };
/* element type of a vector block / precision of the preconditioner */
enum { HH_C128 = 0, HH_C64 = 1 };
/* where a data pointer lives */
enum { HH_HOST = 0, HH_DEVICE = 1 };
/* outer Krylov method (ShiftedLaplacianMultigridSolver.Krylov) */
enum { HH_FGMRES = 0, HH_BICGSTAB = 1 };
/* smoother (MG.relaxType): "Jac" damped Jacobi, "Jac-GMRES" GMRES steps
 * right-preconditioned by damped Jacobi (test/ShiftedLaplacianTest.jl:138) */
enum { HH_RELAX_JAC = 0, HH_RELAX_JAC_GMRES = 1 };
/* per right-hand-side outcome of hh_solve */
enum { HH_CONVERGED = 0, HH_MAXITER = 1, HH_NONFINITE = 2, HH_ZERO_RHS = 3 };

typedef struct hh_op hh_op;
typedef struct hh_solver hh_solver;
typedef struct hh_unet hh_unet;
typedef struct hh_dataset hh_dataset;

/* Multigrid + Krylov settings; mirrors the fields of Multigrid.jl's MGparam
 * that the reference sets through getMGparam (test/ShiftedLaplacianTest.jl:51-64)
 * plus the solver's Krylov/inner (src/ShiftedLaplacianMultigridSolver.jl:8-9). */
typedef struct hh_mg_params {
  int32_t levels;     /* MG.levels (>= 1; every coarsening halves the cells; the
                         coarsest grid, solved by a dense inverse, has at most
                         10000 nodes)                                            */
  int32_t pre;        /* MG.relaxPre: damped-Jacobi sweeps (or Jac-GMRES steps,
                         at most 8) before the correction                      */
  int32_t post;       /* MG.relaxPost                                            */
  double relax;       /* MG.relaxParam: Jacobi weight                            */
  int32_t cycle;      /* MG.cycleType: 'V', 'W' or 'K' (Krylov-accelerated: two
                         minimal-residual steps per coarse visit)              */
  int32_t max_outer;  /* MG.maxOuterIter: FGMRES restarts                        */
  double tol;         /* MG.relativeTol: ||b - H x|| <= tol ||b||                */
  int32_t inner;      /* solver.inner: FGMRES restart length (1..32)             */
  int32_t krylov;     /* HH_FGMRES (inner = restart length) or HH_BICGSTAB       */
  int32_t prec;       /* preconditioner precision: HH_C128 or HH_C64             */
  int32_t learned;    /* reserved: 0 (attach a network with hh_solver_attach_unet) */
  int32_t verbose;    /* solver.verbose                                          */
  int32_t graph;      /* 1: from the second FGMRES restart cycle on, replay each
                         whole cycle (inner steps, x update, true residual) as
                         one captured hipGraph, one host sync per cycle      */
  int32_t autotune;   /* 1: time stencil launch shapes per level at first solve  */
  int32_t fused;      /* 1: LDS-tiled fused smoothing/transfer legs on 2-D levels,
                         z-marching down (pre-smoothing + restricted residual),
                         middle (W: correction + restricted residual) and up
                         legs on 3-D complex64 levels (mirror ghosts, 2 pre /
                         2 post sweeps); 0: one kernel per step */
  int32_t wlevels;    /* with cycle 'W': only levels 0..wlevels-1 apply the double
                         correction, coarser levels a single one (0 = every level) */
  int32_t per_level;  /* 1: sweep counts per level from pre_l/post_l (MG.relaxPre
                         and relaxPost given as functions of the level) */
  int32_t pre_l[16];  /* per_level: sweeps before the correction on level l (0 = finest) */
  int32_t post_l[16]; /* per_level: sweeps after it */
  int32_t relax_type; /* MG.relaxType: HH_RELAX_JAC or HH_RELAX_JAC_GMRES (the
                         fused legs run only with HH_RELAX_JAC) */
  int32_t basis;      /* storage of the FGMRES basis: HH_C128, or HH_C64 (with
                         prec = HH_C64 and the 5/7-point operator: products and
                         sums stay float64, the iterate and the restart residual
                         complex128; halves the Arnoldi passes' traffic) */
} hh_mg_params;

/* Per-call results.  The arrays are caller-owned with nrhs entries, or NULL. */
typedef struct hh_solve_stats {
  int32_t* iters;      /* Krylov iterations per RHS                             */
  int32_t* status;     /* HH_CONVERGED / HH_MAXITER / HH_NONFINITE / HH_ZERO_RHS */
  double* true_res;    /* ||b - H x|| / ||b|| recomputed in complex128          */
  int32_t max_iters;   /* max over RHS                                           */
  int32_t n_converged; /* RHS with status HH_CONVERGED or HH_ZERO_RHS            */
  double setup_time;   /* seconds spent building the hierarchy in this call     */
  double solve_time;   /* seconds in the Krylov solve                            */
  int64_t n_prec;      /* preconditioner applications (lock-step iterations)    */
  int32_t learned_used; /* 1 if the learned correction was applied in this solve */
  int32_t fell_back;    /* 1 if the plain cycle replaced it (trial lost, now or in
                           an earlier solve on this hierarchy, or stagnation)     */
} hh_solve_stats;

const char* hh_last_error(void);
int hh_version(void);

/* Absorbing-layer attenuation profile getABL(n, NeumannAtFirstDim, ABLpad,
 * ABLamp) (src/GetHelmholtz.jl:97-220; the 2-D branch is the reference's
 * active impl = 1, :141-163, the 3-D branch :164-218 with all faces on).
 * out: prod(nnodes) doubles, column-major; loc = HH_HOST | HH_DEVICE. */
int hh_abl(int dim, const int64_t nnodes[3], int neumann_at_first_dim, const int64_t pad[3], double amp,
           double* out, int loc);

/* getABL(n, NeumannAtFirstDim, ABLpad, ABLamp, code) with the reference's face
 * mask `code` (src/GetHelmholtz.jl:97; a length(n) x 2 Bool matrix):
 * faces[2 d] = code[d+1, 1] (low face of axis d), faces[2 d + 1] = code[d+1, 2]
 * (high face), each 0 or 1.  The mask acts in 3-D (:164-218); the 2-D branch
 * (impl = 1, :141-163) does not read it, so 2-D results equal hh_abl's.  A
 * Neumann top still removes the low face of the last axis. */
int hh_abl_faces(int dim, const int64_t nnodes[3], int neumann_at_first_dim, const int64_t pad[3], double amp,
                 const int32_t faces[6], double* out, int loc);

/* Sommerfeld boundary terms getSommerfeldBC(Msh, mNodal, omega, NeumannOnTop,
 * orderNeumannBC) (src/GetHelmholtz.jl:222-247): -i omega (BC / h_d) sqrt(m)
 * summed over the faces of the grid a node lies on (BC = 2 for bc_order 2,
 * 1 for bc_order 1), the top face (low end of the last axis) skipped when
 * neumann_top.  m: prod(nnodes) doubles; out: prod(nnodes) complex128
 * (interleaved), column-major; both host (loc = HH_HOST) or both device. */
int hh_sommerfeld_bc(int dim, const int64_t nnodes[3], const double h[3], const double* m, double omega,
                     int neumann_top, int bc_order, void* out, int loc);

/* Operator H = -Lap_h - w^2 m (1 - i gamma / Re w) + Sommerfeld terms on a
 * nodal grid.  Replaces GetHelmholtzOperator(Msh, m, omega, gamma,
 * NeumannAtFirstDim, Sommerfeld, orderNeumannBC) (src/GetHelmholtz.jl:33-50)
 * and the HelmholtzParam record (src/Helmholtz.jl:13-20): nothing is
 * assembled; m (slowness^2) and gamma (already including the absorbing layer,
 * as returned by the 8-argument GetHelmholtzOperator, src/GetHelmholtz.jl:22-31)
 * are copied to the device.  dim = 2 or 3; nnodes[d] = Msh.n[d] + 1 >= 3;
 * h[d] = Msh.h[d]; bc_order = orderNeumannBC (1 or 2). */
int hh_op_create(int dim, const int64_t nnodes[3], const double h[3], const double* m, const double* gamma,
                 double omega_re, double omega_im, int neumann_top, int sommerfeld, int bc_order, hh_op** out);
/* GetHelmholtzOperatorHO(Msh, m, omega, gamma, NeumannAtFirstDim, Sommerfeld,
 * beta) (src/GetHelmholtz.jl:18-20,54-72; src/PlainNodalLaplacian.jl:54-141):
 * the spread Laplacian G'((1-b)Gs + bG) plus the averaged mass
 * M diag(mass), matrix-free (9-point in 2-D, 19-point in 3-D).  2-D uses
 * beta_lap for both (the reference's scalar beta); 3-D takes the reference's
 * beta = [beta_lap; beta_mass].  Applies in complex128 (hh_op_apply); as the
 * operator of hh_solve it is the outer operator and the preconditioner's
 * hierarchy uses the natural-Neumann (BC = 1) 5/7-point stencil, which is
 * this operator at beta = 1.  Adjoint solves (hh_solve, do_transpose = 1)
 * apply its transpose Lap + diag(mass) M. */
int hh_op_create_ho(int dim, const int64_t nnodes[3], const double h[3], const double* m, const double* gamma,
                    double omega_re, double omega_im, int neumann_top, int sommerfeld, double beta_lap,
                    double beta_mass, hh_op** out);
int hh_op_destroy(hh_op* op);

/* Y = (H + i*shift*w^2*diag(m)) X for nrhs columns.  shift = 0 gives H·x, the
 * product ParSpMatVec's SpMatMul forms inside getHelmholtzFun
 * (src/GetHelmholtz.jl:85-95); shift = s gives the operator of
 * getShiftedHelmholtzParam(p, s) (src/Helmholtz.jl:32-34), i.e. H +
 * GetHelmholtzShiftOP(m, w, s) for real w (src/GetHelmholtz.jl:81-83).
 * prec = HH_C128 | HH_C64 (element type of X and Y), loc = HH_HOST | HH_DEVICE. */
int hh_op_apply(hh_op* op, double shift, const void* X, void* Y, int64_t nrhs, int prec, int loc);

/* hh_op_apply on host blocks X, Y that allocates nothing: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).  X
 * and Y are staged through a fixed scratch array of 64 MiB in device memory
 * (32 MiB in, 32 MiB out), which each GPU holds from the library's load to
 * the end of the process; a call with N * nrhs * (16 | 8) bytes above
 * 32 MiB (33,554,432) is refused with HH_ERR_ARG.  The operator itself is
 * created beforehand (hh_op_create allocates).  Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU race.  Results equal
 * hh_op_apply's bit for bit. */
int hh_op_apply_noalloc(hh_op* op, double shift, const void* X, void* Y, int64_t nrhs, int prec);

/* Shifted-Laplacian multigrid solver bound to op (which must outlive it).
 * Replaces getShiftedLaplacianMultigridSolver(helmParam, MG, shift, Krylov,
 * inner, verbose) (src/ShiftedLaplacianMultigridSolver.jl:24-30).
 * shift_per_level: p->levels values (the reference broadcasts a scalar,
 * :28-30).  The hierarchy is built lazily by the first hh_solve, as the
 * reference's MGsetup is (:50-66). */
int hh_solver_create(hh_op* op, const hh_mg_params* p, const double* shift_per_level, hh_solver** out);
/* copySolver (src/ShiftedLaplacianMultigridSolver.jl:18-22): same settings, no hierarchy. */
int hh_solver_copy(const hh_solver* s, hh_solver** out);
/* clear! (src/ShiftedLaplacianMultigridSolver.jl:105-109): drops the hierarchy. */
int hh_solver_clear(hh_solver* s);
int hh_solver_destroy(hh_solver* s);

/* solveLinearSystem(ShiftedHT, B, param, doTranspose)
 * (src/ShiftedLaplacianMultigridSolver.jl:33-102): X = H^-1 B column by
 * column to ||b - H x|| <= tol ||b|| with FGMRES(inner) preconditioned by one
 * shifted-Laplacian MG cycle per iteration.  B and X are complex128, N x nrhs.
 * An all-zero B returns X = 0 (:40-43).  krylov = HH_BICGSTAB runs
 * right-preconditioned BiCGSTAB instead (:91-93; two cycles per iteration, at
 * most max_outer iterations).  do_transpose = 1 solves the adjoint system
 * H^H X = B (:68-80), through the mirror-ghost similarity
 * H^H = W conj(H) W^-1 (W: nodal trapezoid weights) with the same hierarchy
 * (for the spread operator of hh_op_create_ho: H^T conj(X) = conj(B), its
 * transpose applied by the same kernel); stats then report the adjoint
 * system's true residuals.
 * stats may be NULL. */
int hh_solve(hh_solver* s, const void* B, void* X, int64_t nrhs, int do_transpose, int loc, hh_solve_stats* stats);

/* One preconditioner application Z = M^-1 V (V complex128, Z in the
 * preconditioner's precision; N x nrhs).  Builds the hierarchy if needed.
 * Inspection/test hook with no counterpart in the reference. */
int hh_solver_precondition(hh_solver* s, const void* V, void* Z, int64_t nrhs, int loc);

/* Per-kernel timing inside hh_solve (instrumentation with no counterpart in
 * the reference): every later launch of `kernel` on grid level `level` is
 * bracketed by HIP events on the solver's stream, and its elapsed time,
 * launch count and algorithmic bytes (the bytes each launch must move for the
 * RHS still iterating) accumulate until the next hh_solver_time_kernel call.
 * HH_KT_NONE switches it off.  Graph replay is not used while timing is on.
 * Algorithmic bytes per launch (s = bytes per preconditioner value, N / Nc =
 * nodes of the level / the next coarser level, a = active RHS; the shared
 * diagonal and Jacobi factor count once per launch):
 *   DOWN   a (2 N s + Nc s) + 2 N s  (a (N s + Nc s) + 2 N s on W levels,
 *          where the smoothed iterate is not stored)
 *   MID    a (2 N s + 2 Nc s) + 2 N s  (W levels: b and e_c read, the
 *          smoothed iterate recomputed from b, x' and b_c written)
 *   UP     a (3 N s + Nc s) + N s  (x, b, e_c read, x'' written; the
 *          sweeps need the Jacobi factor, not the diagonal)
 *   APPLY  a N (s_z + 16) + 16 N
 *   DOTS   16 a N (j + 2)              UPDATE 16 a N (j + 3)  (Arnoldi step j)
 * With prec = HH_C64 on the 5/7-point operator the passes form H z from the
 * complex64 z (v = 16 or 8 B per basis value, t = 8 B for a staged copy with
 * the complex128 basis, else 0):
 *   DOTS   a N (v (j + 1) + 8) + 16 N  UPDATE a N (v (j + 2) + 8 + t) + 16 N
 * and APPLY (a separate H z launch) exists only with the complex128 basis:
 * timing it switches that solve to the stored-w passes, with the same
 * iterates; with basis = HH_C64 it is refused. */
enum {
  HH_KT_NONE = 0, HH_KT_DOWN = 1, HH_KT_MID = 2, HH_KT_UP = 3, HH_KT_APPLY = 4, HH_KT_DOTS = 5, HH_KT_UPDATE = 6
};
int hh_solver_time_kernel(hh_solver* s, int kernel, int level);
int hh_solver_kernel_time(hh_solver* s, double* total_ms, int64_t* launches, double* alg_bytes);

/* Number of grid levels the solver's hierarchy has (0 before the first solve)
 * and the node counts of level l. */
int hh_solver_levels(const hh_solver* s, int* levels, int64_t nnodes_of_level[][3], int max_levels);

/* Device memory for callers that keep their blocks resident in HBM (no
 * counterpart in the reference, whose arrays live in host memory).  kind:
 * 0 host->device, 1 device->host, 2 device->device. */
int hh_device_alloc(int64_t bytes, void** ptr);
int hh_device_free(void* ptr);
int hh_memcpy(void* dst, const void* src, int64_t bytes, int kind);

/* Measurement hook (no counterpart in the reference): launches one operator
 * kernel `reps` times on the operator's stream between two HIP events and
 * returns the average duration per launch.  kernel: 0 Y = A X, 1 Y = B - A X,
 * 2 one damped-Jacobi sweep Y = X + 0.8 D^-1 (B - A X); A = H + i shift w^2 m.
 * X, B, Y are device blocks of nrhs columns in precision prec. */
int hh_op_time_kernel(hh_op* op, int kernel, double shift, const void* X, const void* B, void* Y, int64_t nrhs,
                      int prec, int reps, double* avg_ms);

/* ---- Learned coarse-grid correction (no counterpart in the reference) ----
 * A 4-level U-Net (16/32/64/64 channels, 3x3 convolutions, ReLU, 1x1 head)
 * in BF16 with FP32 accumulation on MFMA.  Input features per node of the
 * learned level: Re r / rms(r), Im r / rms(r), Re(w)^2 m h_x^2, gamma / Re(w);
 * output: the correction e = rms(r) h_x^2 (y0 + i y1) for A_l e = r.
 * Parameters are fp32 (flat layout: per conv layer W[9][co][ci] then bias,
 * then the head's W[2][16] and bias[2]). */
int64_t hh_unet_num_params(void);
int hh_unet_create(uint64_t seed, hh_unet** out);
int hh_unet_destroy(hh_unet* u);
int hh_unet_get_params(const hh_unet* u, float* out);
int hh_unet_set_params(hh_unet* u, const float* in);
/* feats: host [nimg][nx*ny][4] fp32 (rounded to BF16 on the device); y: host [nimg][nx*ny][2] */
int hh_unet_forward(hh_unet* u, int nimg, int nx, int ny, const float* feats, float* y);
/* mean-squared-error loss against target [nimg][nx*ny][2] and its gradient w.r.t. the flat parameters */
int hh_unet_loss_grad(hh_unet* u, int nimg, int nx, int ny, const float* feats, const float* target, float* loss,
                      float* grad);
/* one Adam step with `grad` (host, flat; NULL = the last computed gradient) */
int hh_unet_adam_step(hh_unet* u, const float* grad, float lr, float beta1, float beta2, float eps);

/* Training pairs on a fixed 2-D grid of nx x ny nodes (device resident). */
int hh_dataset_create(int nx, int ny, int64_t capacity, hh_dataset** out);
int hh_dataset_destroy(hh_dataset* d);
int hh_dataset_size(const hh_dataset* d, int64_t* n);
/* Record the right-hand sides that reach grid level `level` (first visit per
 * preconditioner application) during FGMRES iterations iter_lo..iter_hi-1
 * (counted across restarts, while every RHS is still active) of the solver's
 * next hh_solve calls (complex64 preconditioner), up to max_samples. */
int hh_solver_capture_begin(hh_solver* s, int level, int64_t max_samples, int iter_lo, int iter_hi);
int hh_solver_capture_count(const hh_solver* s, int64_t* n);
/* Append the captured residuals r with targets e = A_level^-1 r and the
 * level's features.  Targets are solved to 1e-6 by FGMRES(10) on the level's
 * shifted operator (W(2,2) preconditioner with a further 0.5 shift), at most
 * max(4, sweeps) restarts. */
int hh_dataset_add_from_solver(hh_dataset* d, hh_solver* s, int sweeps);
/* Adam training on the dataset (mean-squared error on normalised targets). */
int hh_unet_train(hh_unet* u, hh_dataset* d, int epochs, int batch, float lr, uint64_t seed, float* loss_per_epoch);
/* mean relative error ||e_net - e|| / ||e|| over samples first..first+count-1 */
int hh_unet_eval(hh_unet* u, hh_dataset* d, int64_t first, int64_t count, float* rel_err);
/* Use the network as the coarse-grid correction at grid level `level` (its
 * output replaces the cycle below that level) inside the FGMRES solve.  The
 * true residual decides: restart cycle 0 runs the network, cycle 1 the plain
 * cycle, and the solve keeps the network only if its cycle reduced the
 * residual (geometric mean over RHS) below 0.8x the plain cycle's reduction;
 * afterwards the plain cycle also takes over if a restart cycle stops
 * reducing every residual by 10% or produces non-finite values.  The verdict
 * is kept for later solves on the same hierarchy (they skip the trial) until
 * the network is re-attached or the hierarchy is rebuilt.
 * hh_solve_stats.fell_back reports the plain cycle was used.  net = NULL detaches. */
int hh_solver_attach_unet(hh_solver* s, hh_unet* net, int level);
/* How an attached network is used: HH_LEARN_TRIAL (default) is the trial
 * described above; HH_LEARN_ALWAYS runs the network in every restart cycle of
 * every solve (no trial, no stagnation switch; only non-finite results bring
 * in the plain cycle) — the ablation of the learned correction.  Resets the
 * kept verdict. */
enum { HH_LEARN_TRIAL = 0, HH_LEARN_ALWAYS = 1 };
int hh_solver_learned_policy(hh_solver* s, int policy);
/* Inspection hook (no counterpart in the reference): device addresses of the
 * solver's own network workspace — its first activation buffer and its input
 * staging block — or NULL where the network has not run on this solver yet.
 * Solvers holding one network share only its read-only weights, never these. */
int hh_solver_learned_workspace(const hh_solver* s, const void** activations, const void** staging);

/* ---- Uneven stream of independent problems (no counterpart in the reference) ---- */
typedef struct hh_problem {
  int32_t dim;                 /* 2 or 3 */
  int64_t nnodes[3];
  double h[3];
  const double* m;             /* host, N */
  const double* gamma;         /* host, N (absorbing layer included) */
  double omega_re, omega_im;
  int32_t neumann_top, sommerfeld, bc_order;
  hh_mg_params params;
  const double* shift_per_level;
  const void* B;               /* complex128 N x nrhs */
  void* X;
  int64_t nrhs;
  int32_t loc;                 /* HH_HOST / HH_DEVICE */
  int32_t* iters;              /* optional per-RHS outputs (nrhs each) */
  int32_t* status;
  double* true_res;
  hh_unet* net;                /* optional learned correction */
  int32_t learn_level;
} hh_problem;

typedef struct hh_problem_result {
  int32_t code;                /* HH_OK or the problem's error code */
  int32_t max_iters;
  int32_t n_converged;
  int32_t fell_back;
  double seconds;              /* wall time of create + solve + destroy */
} hh_problem_result;

/* Solve every problem (largest N * nrhs first) with `nworkers` host threads,
 * each problem on its own stream; the call returns when all are done.  A
 * handle per worker thread; a failing problem reports its code and the call
 * returns the first error after the others finished. */
int hh_stream_solve(int nprob, const hh_problem* probs, int nworkers, hh_problem_result* results);

#ifdef __cplusplus
}
#endif
#endif /* HELMHOLTZ_AMD_H */
