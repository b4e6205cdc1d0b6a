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
/* per right-hand-side outcome of hh_solve */
enum { HH_CONVERGED = 0, HH_MAXITER = 1, HH_NONFINITE = 2, HH_ZERO_RHS = 3 };

typedef struct hh_op hh_op;
typedef struct hh_solver hh_solver;

/* Multigrid + Krylov settings; mirrors the fields of Multigrid.jl's MGparam
 * that the reference sets through getMGparam (test/ShiftedLaplacianTest.jl:51-64)
 * plus the solver's Krylov/inner (src/ShiftedLaplacianMultigridSolver.jl:8-9). */
typedef struct hh_mg_params {
  int32_t levels;     /* MG.levels (>= 1; every coarsening halves the cells)     */
  int32_t pre;        /* MG.relaxPre: damped-Jacobi sweeps before the correction */
  int32_t post;       /* MG.relaxPost                                            */
  double relax;       /* MG.relaxParam: Jacobi weight                            */
  int32_t cycle;      /* MG.cycleType: 'V' or 'W'                                */
  int32_t max_outer;  /* MG.maxOuterIter: FGMRES restarts                        */
  double tol;         /* MG.relativeTol: ||b - H x|| <= tol ||b||                */
  int32_t inner;      /* solver.inner: FGMRES restart length (1..32)             */
  int32_t krylov;     /* HH_FGMRES                                               */
  int32_t prec;       /* preconditioner precision: HH_C128 or HH_C64             */
  int32_t learned;    /* reserved: 0                                             */
  int32_t verbose;    /* solver.verbose                                          */
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
} hh_solve_stats;

const char* hh_last_error(void);
int hh_version(void);

/* Absorbing-layer attenuation profile getABL(n, NeumannAtFirstDim, ABLpad,
 * ABLamp) (src/GetHelmholtz.jl:97-220; the 2-D branch is the reference's
 * active impl = 1, :141-163, the 3-D branch :164-218 with all faces on).
 * out: prod(nnodes) doubles, column-major; loc = HH_HOST | HH_DEVICE. */
int hh_abl(int dim, const int64_t nnodes[3], int neumann_at_first_dim, const int64_t pad[3], double amp,
           double* out, int loc);

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
int hh_op_destroy(hh_op* op);

/* Y = (H + i*shift*w^2*diag(m)) X for nrhs columns.  shift = 0 gives H·x, the
 * product ParSpMatVec's SpMatMul forms inside getHelmholtzFun
 * (src/GetHelmholtz.jl:85-95); shift = s gives the operator of
 * getShiftedHelmholtzParam(p, s) (src/Helmholtz.jl:32-34), i.e. H +
 * GetHelmholtzShiftOP(m, w, s) for real w (src/GetHelmholtz.jl:81-83).
 * prec = HH_C128 | HH_C64 (element type of X and Y), loc = HH_HOST | HH_DEVICE. */
int hh_op_apply(hh_op* op, double shift, const void* X, void* Y, int64_t nrhs, int prec, int loc);

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
 * An all-zero B returns X = 0 (:40-43).  do_transpose must be 0.
 * stats may be NULL. */
int hh_solve(hh_solver* s, const void* B, void* X, int64_t nrhs, int do_transpose, int loc, hh_solve_stats* stats);

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

#ifdef __cplusplus
}
#endif
#endif /* HELMHOLTZ_AMD_H */
