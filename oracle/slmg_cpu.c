/* CPU restatement of the plain shifted-Laplacian multigrid FGMRES solve, in C
 * with OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY: the timed CPU baseline beside the GPU solver and a
 * second, independent check of the NumPy oracle (oracle/slmg_ref.py).  The
 * product path never links or loads it.
 *
 * It is the reference's plain method on assembled sparse matrices, as the
 * reference runs it on the CPU:
 *   - H assembled in CSR exactly as GetHelmholtzOperator does
 *     (src/GetHelmholtz.jl:33-50): Laplacian = sum over axes of the 1-D
 *     second difference dxxMat (src/PlainNodalLaplacian.jl:18-30, rows
 *     [BC,-BC]/h^2, [-1,2,-1]/h^2, [-BC,BC]/h^2) plus the diagonal
 *     -w^2 m (1 - i gamma / Re w) + i Re(w) sqrt(m) sum_faces 2/h
 *     (getSommerfeldBC, src/GetHelmholtz.jl:222-247, always BC = 2, top face
 *     skipped with NeumannOnTop);
 *   - rediscretised shifted hierarchy (src/ShiftedLaplacianMultigridSolver.jl:51-62):
 *     gamma += shift[0] Re w on the fine level, m and gamma restricted by
 *     full weighting with mirror ghosts, gamma += (shift[l+1]-shift[l]) Re w;
 *   - damped Jacobi by CSR SpMV, full-weighting restriction, bi/trilinear
 *     prolongation, dense LU on the coarsest level;
 *   - V or W cycle (W: the coarse correction applied twice per visit, each
 *     from the residual recomputed on that level, on levels < wlevels);
 *   - right-preconditioned FGMRES(inner) with modified Gram-Schmidt and
 *     restarts, stopping on the Arnoldi estimate ||r|| <= tol ||b||
 *     (the control flow of solveLinearSystem, :33-102).
 * Every loop over nodes is an OpenMP parallel loop (the reference's
 * numCores threads in SpMatMul, src/GetHelmholtz.jl:90-91).
 */
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef double _Complex cplx;

typedef struct {
  int64_t n;
  int64_t* rp;
  int32_t* ci;
  cplx* v;
} csr;

typedef struct {
  int dim;
  int64_t nn[3];
  double h[3];
  int64_t N;
  double *m, *g;
  csr A;
  cplx* dinv;
  cplx *x, *b, *r;   /* level work vectors (x, b of the cycle; r residual) */
  cplx* lu;          /* coarsest: dense LU, row-major */
  int* piv;
} level;

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* ---------------------------------------------------------------- assembly */

/* CSR of the operator on one level; src/GetHelmholtz.jl:33-50. */
static void assemble(level* L, double wre, double wim, int neumann, int sommerfeld, int bc) {
  const int dim = L->dim;
  const int64_t N = L->N;
  const int64_t n0 = L->nn[0], n1 = L->nn[1], n2 = dim == 3 ? L->nn[2] : 1;
  const int64_t stride[3] = {1, n0, n0 * n1};
  const cplx w = wre + I * wim, w2 = w * w;
  L->A.n = N;
  L->A.rp = malloc(sizeof(int64_t) * (N + 1));
  L->A.ci = malloc(sizeof(int32_t) * N * (2 * dim + 1));
  L->A.v = malloc(sizeof(cplx) * N * (2 * dim + 1));
  L->dinv = malloc(sizeof(cplx) * N);
  /* every row holds the diagonal plus one or two neighbours per axis */
  L->A.rp[0] = 0;
  for (int64_t p = 0; p < N; ++p) {
    int64_t idx[3] = {p % n0, (p / n0) % n1, dim == 3 ? p / (n0 * n1) : 0};
    int cnt = 1;
    for (int d = 0; d < dim; ++d) {
      if (L->nn[d] < 2) continue;
      cnt += (idx[d] == 0 || idx[d] == L->nn[d] - 1) ? 1 : 2;
    }
    L->A.rp[p + 1] = L->A.rp[p] + cnt;
  }
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < N; ++p) {
    int64_t idx[3] = {p % n0, (p / n0) % n1, dim == 3 ? p / (n0 * n1) : 0};
    (void)n2;
    int64_t q = L->A.rp[p];
    cplx diag = -w2 * L->m[p] * (1.0 - I * L->g[p] / wre);
    if (sommerfeld) {
      const double sm = sqrt(L->m[p]);
      for (int d = 0; d < dim; ++d) {
        const double c = 2.0 / L->h[d];
        const int top = d == dim - 1;
        if (idx[d] == 0 && !(top && neumann)) diag += I * wre * c * sm;
        if (idx[d] == L->nn[d] - 1) diag += I * wre * c * sm;
      }
    }
    /* neighbours in increasing column order: axes from slowest to fastest
       below the diagonal, fastest to slowest above it */
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
      const double ih2 = 1.0 / (L->h[d] * L->h[d]);
      if (L->nn[d] < 2) continue;
      if (idx[d] == 0) {
        diag += bc * ih2;
        hi[d] = -bc * ih2;
      } else if (idx[d] == L->nn[d] - 1) {
        diag += bc * ih2;
        lo[d] = -bc * ih2;
      } else {
        diag += 2.0 * ih2;
        lo[d] = -ih2;
        hi[d] = -ih2;
      }
    }
    for (int d = dim - 1; d >= 0; --d)
      if (lo[d] != 0.0) L->A.ci[q] = (int32_t)(p - stride[d]), L->A.v[q++] = lo[d];
    L->A.ci[q] = (int32_t)p, L->A.v[q++] = diag;
    for (int d = 0; d < dim; ++d)
      if (hi[d] != 0.0) L->A.ci[q] = (int32_t)(p + stride[d]), L->A.v[q++] = hi[d];
    L->dinv[p] = 1.0 / diag;
  }
}

static void spmv(const csr* A, const cplx* x, cplx* y) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < A->n; ++p) {
    cplx s = 0;
    for (int64_t q = A->rp[p]; q < A->rp[p + 1]; ++q) s += A->v[q] * x[A->ci[q]];
    y[p] = s;
  }
}

/* ------------------------------------------------------------- transfers */

/* 1-D full weighting along axis d with mirror ghosts:
   c[I] = (f[2I-1] + 2 f[2I] + f[2I+1]) / 4, f[-1] := f[1], f[nf] := f[nf-2]. */
static void restrict_axis(const cplx* f, cplx* c, const int64_t* nf, int d, int dim) {
  int64_t nc[3] = {nf[0], nf[1], dim == 3 ? nf[2] : 1};
  int64_t fn[3] = {nf[0], nf[1], dim == 3 ? nf[2] : 1};
  nc[d] = (nf[d] - 1) / 2 + 1;
  const int64_t total = nc[0] * nc[1] * nc[2];
  const int64_t fs[3] = {1, fn[0], fn[0] * fn[1]};
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < total; ++p) {
    int64_t i[3] = {p % nc[0], (p / nc[0]) % nc[1], p / (nc[0] * nc[1])};
    int64_t ff[3] = {i[0], i[1], i[2]};
    ff[d] = 2 * i[d];
    const int64_t base = ff[0] * fs[0] + ff[1] * fs[1] + ff[2] * fs[2];
    const int64_t l = ff[d] > 0 ? -fs[d] : fs[d];
    const int64_t r = ff[d] < fn[d] - 1 ? fs[d] : -fs[d];
    c[p] = 0.5 * f[base] + 0.25 * f[base + l] + 0.25 * f[base + r];
  }
}

/* 1-D linear interpolation along axis d: f[2I] = c[I], f[2I+1] = (c[I]+c[I+1])/2. */
static void prolong_axis(const cplx* c, cplx* f, const int64_t* ncin, int d, int dim) {
  int64_t nc[3] = {ncin[0], ncin[1], dim == 3 ? ncin[2] : 1};
  int64_t fn[3] = {nc[0], nc[1], nc[2]};
  fn[d] = 2 * (nc[d] - 1) + 1;
  const int64_t total = fn[0] * fn[1] * fn[2];
  const int64_t cs[3] = {1, nc[0], nc[0] * nc[1]};
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < total; ++p) {
    int64_t i[3] = {p % fn[0], (p / fn[0]) % fn[1], p / (fn[0] * fn[1])};
    int64_t cc[3] = {i[0], i[1], i[2]};
    cc[d] = i[d] / 2;
    const int64_t base = cc[0] * cs[0] + cc[1] * cs[1] + cc[2] * cs[2];
    f[p] = (i[d] & 1) ? 0.5 * (c[base] + c[base + cs[d]]) : c[base];
  }
}

/* fine -> coarse full weighting over every axis (tensor product) */
static void restrict_full(const cplx* f, cplx* c, const int64_t* nf, int dim, cplx* t1, cplx* t2) {
  int64_t n[3] = {nf[0], nf[1], dim == 3 ? nf[2] : 1};
  const cplx* src = f;
  for (int d = 0; d < dim; ++d) {
    cplx* dst = d == dim - 1 ? c : (d % 2 == 0 ? t1 : t2);
    restrict_axis(src, dst, n, d, dim);
    n[d] = (n[d] - 1) / 2 + 1;
    src = dst;
  }
}

/* coarse -> fine interpolation, added to xf */
static void prolong_add(const cplx* c, cplx* xf, const int64_t* ncin, int dim, cplx* t1, cplx* t2, int64_t Nf) {
  int64_t n[3] = {ncin[0], ncin[1], dim == 3 ? ncin[2] : 1};
  const cplx* src = c;
  for (int d = 0; d < dim; ++d) {
    cplx* dst = d % 2 == 0 ? t1 : t2;
    prolong_axis(src, dst, n, d, dim);
    n[d] = 2 * (n[d] - 1) + 1;
    src = dst;
  }
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < Nf; ++p) xf[p] += src[p];
}

/* ------------------------------------------------------------- coarsest */

static int dense_lu(level* L) {
  const int64_t n = L->N;
  L->lu = calloc((size_t)(n * n), sizeof(cplx));
  L->piv = malloc(sizeof(int) * n);
  if (!L->lu || !L->piv) return -1;
  for (int64_t p = 0; p < n; ++p)
    for (int64_t q = L->A.rp[p]; q < L->A.rp[p + 1]; ++q) L->lu[p * n + L->A.ci[q]] += L->A.v[q];
  cplx* a = L->lu;
  for (int64_t k = 0; k < n; ++k) {
    int64_t pv = k;
    double best = cabs(a[k * n + k]);
    for (int64_t i = k + 1; i < n; ++i)
      if (cabs(a[i * n + k]) > best) best = cabs(a[i * n + k]), pv = i;
    L->piv[k] = (int)pv;
    if (best == 0.0) return -2;
    if (pv != k)
      for (int64_t j = 0; j < n; ++j) {
        cplx t = a[k * n + j];
        a[k * n + j] = a[pv * n + j];
        a[pv * n + j] = t;
      }
    const cplx inv = 1.0 / a[k * n + k];
#pragma omp parallel for schedule(static)
    for (int64_t i = k + 1; i < n; ++i) {
      const cplx f = a[i * n + k] * inv;
      a[i * n + k] = f;
      for (int64_t j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
    }
  }
  return 0;
}

static void dense_solve(const level* L, const cplx* b, cplx* x) {
  const int64_t n = L->N;
  const cplx* a = L->lu;
  memcpy(x, b, sizeof(cplx) * n);
  for (int64_t k = 0; k < n; ++k) {
    const int64_t pv = L->piv[k];
    if (pv != k) {
      cplx t = x[k];
      x[k] = x[pv];
      x[pv] = t;
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    cplx s = x[i];
    for (int64_t j = 0; j < i; ++j) s -= a[i * n + j] * x[j];
    x[i] = s;
  }
  for (int64_t i = n - 1; i >= 0; --i) {
    cplx s = x[i];
    for (int64_t j = i + 1; j < n; ++j) s -= a[i * n + j] * x[j];
    x[i] = s / a[i * n + i];
  }
}

/* ---------------------------------------------------------------- cycle */

typedef struct {
  int nlev, pre, post, cycle_w, wlevels;
  double relax;
  level* L;
  cplx *t1, *t2;  /* transfer temporaries, fine-level size */
} hierarchy;

static void jacobi(const level* L, cplx* x, const cplx* b, cplx* r, double w, int sweeps) {
  for (int s = 0; s < sweeps; ++s) {
    spmv(&L->A, x, r);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < L->N; ++p) x[p] += w * L->dinv[p] * (b[p] - r[p]);
  }
}

/* x_l = cycle(b_l) from x_l = 0; slmg_ref.vcycle */
static void cycle(hierarchy* Hh, int l) {
  level* L = &Hh->L[l];
  if (l == Hh->nlev - 1) {
    dense_solve(L, L->b, L->x);
    return;
  }
  memset(L->x, 0, sizeof(cplx) * L->N);
  jacobi(L, L->x, L->b, L->r, Hh->relax, Hh->pre);
  const int ncorr = (Hh->cycle_w && (Hh->wlevels <= 0 || l < Hh->wlevels)) ? 2 : 1;
  level* C = &Hh->L[l + 1];
  for (int k = 0; k < ncorr; ++k) {
    spmv(&L->A, L->x, L->r);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < L->N; ++p) L->r[p] = L->b[p] - L->r[p];
    restrict_full(L->r, C->b, L->nn, L->dim, Hh->t1, Hh->t2);
    cycle(Hh, l + 1);
    prolong_add(C->x, L->x, C->nn, L->dim, Hh->t1, Hh->t2, L->N);
  }
  jacobi(L, L->x, L->b, L->r, Hh->relax, Hh->post);
}

static void free_hierarchy(hierarchy* Hh) {
  if (!Hh->L) return;
  for (int l = 0; l < Hh->nlev; ++l) {
    level* L = &Hh->L[l];
    free(L->m), free(L->g), free(L->A.rp), free(L->A.ci), free(L->A.v), free(L->dinv);
    free(L->x), free(L->b), free(L->r), free(L->lu), free(L->piv);
  }
  free(Hh->L), free(Hh->t1), free(Hh->t2);
}

static int build(hierarchy* Hh, int dim, const int64_t* nn, const double* h, const double* m, const double* gamma,
                 double wre, double wim, int neumann, int sommerfeld, int bc, int levels, const double* shift) {
  Hh->L = calloc(levels, sizeof(level));
  Hh->nlev = levels;
  int64_t N0 = 1;
  for (int d = 0; d < dim; ++d) N0 *= nn[d];
  Hh->t1 = malloc(sizeof(cplx) * N0);
  Hh->t2 = malloc(sizeof(cplx) * N0);
  cplx* ctmp = malloc(sizeof(cplx) * N0);
  cplx* cout = malloc(sizeof(cplx) * N0);
  for (int l = 0; l < levels; ++l) {
    level* L = &Hh->L[l];
    L->dim = dim;
    L->N = 1;
    for (int d = 0; d < dim; ++d) {
      L->nn[d] = l == 0 ? nn[d] : (Hh->L[l - 1].nn[d] - 1) / 2 + 1;
      L->h[d] = l == 0 ? h[d] : 2.0 * Hh->L[l - 1].h[d];
      L->N *= L->nn[d];
    }
    L->m = malloc(sizeof(double) * L->N);
    L->g = malloc(sizeof(double) * L->N);
    if (l == 0) {
      for (int64_t p = 0; p < L->N; ++p) L->m[p] = m[p], L->g[p] = gamma[p] + shift[0] * wre;
    } else {
      level* F = &Hh->L[l - 1];
      for (int d = 0; d < dim; ++d)
        if ((F->nn[d] - 1) % 2) return -3;  /* cells must be even to coarsen */
      for (int pass = 0; pass < 2; ++pass) {
        const double* src = pass ? F->g : F->m;
        for (int64_t p = 0; p < F->N; ++p) ctmp[p] = src[p];
        restrict_full(ctmp, cout, F->nn, dim, Hh->t1, Hh->t2);
        for (int64_t p = 0; p < L->N; ++p) {
          if (pass)
            L->g[p] = creal(cout[p]) + (shift[l] - shift[l - 1]) * wre;
          else
            L->m[p] = creal(cout[p]);
        }
      }
    }
    assemble(L, wre, wim, neumann, sommerfeld, bc);
    L->x = calloc(L->N, sizeof(cplx));
    L->b = calloc(L->N, sizeof(cplx));
    L->r = calloc(L->N, sizeof(cplx));
  }
  free(ctmp), free(cout);
  return dense_lu(&Hh->L[levels - 1]);
}

/* --------------------------------------------------------------- FGMRES */

static cplx dotc(const cplx* a, const cplx* b, int64_t n) {
  double re = 0, im = 0;
#pragma omp parallel for reduction(+ : re, im) schedule(static)
  for (int64_t p = 0; p < n; ++p) {
    const cplx t = conj(a[p]) * b[p];
    re += creal(t), im += cimag(t);
  }
  return re + I * im;
}

static double nrm2(const cplx* a, int64_t n) { return sqrt(creal(dotc(a, a, n))); }

/* slmg_ref.fgmres: right-preconditioned FGMRES(inner), MGS, restarts */
static int fgmres(hierarchy* Hh, const level* F, const cplx* b, cplx* x, int inner, int max_outer, double tol,
                  int* iters) {
  const int64_t n = F->N;
  const int K = inner;
  cplx* V = malloc(sizeof(cplx) * n * (K + 1));
  cplx* Z = malloc(sizeof(cplx) * n * K);
  cplx* w = malloc(sizeof(cplx) * n);
  cplx* Hm = malloc(sizeof(cplx) * (K + 1) * K);
  cplx* cs = malloc(sizeof(cplx) * K);
  cplx* sn = malloc(sizeof(cplx) * K);
  cplx* gv = malloc(sizeof(cplx) * (K + 1));
  cplx* y = malloc(sizeof(cplx) * K);
  int total = 0;
  memset(x, 0, sizeof(cplx) * n);
  const double bn = nrm2(b, n);
  if (bn == 0.0) {
    *iters = 0;
    goto done;
  }
  for (int outer = 0; outer < max_outer; ++outer) {
    spmv(&F->A, x, w);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) w[p] = b[p] - w[p];
    const double beta = nrm2(w, n);
    if (beta <= tol * bn) break;
    memset(Hm, 0, sizeof(cplx) * (K + 1) * K);
    memset(gv, 0, sizeof(cplx) * (K + 1));
    gv[0] = beta;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) V[p] = w[p] / beta;
    int k = 0;
    for (int j = 0; j < K; ++j) {
      level* L0 = &Hh->L[0];
      memcpy(L0->b, V + (int64_t)j * n, sizeof(cplx) * n);
      cycle(Hh, 0);
      memcpy(Z + (int64_t)j * n, L0->x, sizeof(cplx) * n);
      spmv(&F->A, Z + (int64_t)j * n, w);
#define HM(i, jj) Hm[(int64_t)(i) * K + (jj)]
      for (int i = 0; i <= j; ++i) {
        const cplx* vi = V + (int64_t)i * n;
        const cplx hij = dotc(vi, w, n);
        HM(i, j) = hij;
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < n; ++p) w[p] -= hij * vi[p];
      }
      const double hn = nrm2(w, n);
      HM(j + 1, j) = hn;
      if (hn != 0.0) {
        cplx* vn = V + (int64_t)(j + 1) * n;
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < n; ++p) vn[p] = w[p] / hn;
      }
      for (int i = 0; i < j; ++i) {
        const cplx t = cs[i] * HM(i, j) + sn[i] * HM(i + 1, j);
        HM(i + 1, j) = -conj(sn[i]) * HM(i, j) + conj(cs[i]) * HM(i + 1, j);
        HM(i, j) = t;
      }
      const cplx a = HM(j, j), bb = HM(j + 1, j);
      const double den = sqrt(creal(a) * creal(a) + cimag(a) * cimag(a) + creal(bb) * creal(bb) +
                              cimag(bb) * cimag(bb));
      cs[j] = den != 0.0 ? conj(a) / den : 1.0;
      sn[j] = den != 0.0 ? conj(bb) / den : 0.0;
      HM(j, j) = den;
      HM(j + 1, j) = 0.0;
      gv[j + 1] = -conj(sn[j]) * gv[j];
      gv[j] = cs[j] * gv[j];
      k = j + 1;
      ++total;
      if (cabs(gv[j + 1]) <= tol * bn) break;
    }
    for (int i = k - 1; i >= 0; --i) {
      cplx s = gv[i];
      for (int jj = i + 1; jj < k; ++jj) s -= HM(i, jj) * y[jj];
      y[i] = s / HM(i, i);
    }
#undef HM
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
      cplx s = 0;
      for (int i = 0; i < k; ++i) s += Z[(int64_t)i * n + p] * y[i];
      x[p] += s;
    }
    if (cabs(gv[k]) <= tol * bn) break;
  }
  *iters = total;
done:
  free(V), free(Z), free(w), free(Hm), free(cs), free(sn), free(gv), free(y);
  return 0;
}

/* ------------------------------------------------------------------ ABI */

int slmg_cpu_threads(void) { return omp_get_max_threads(); }

/* y = H x on the assembled operator (pins the CSR assembly) */
int slmg_cpu_apply(int dim, const int64_t* nn, const double* h, const double* m, const double* gamma, double wre,
                   double wim, int neumann, int sommerfeld, int bc, const double* x, double* y) {
  level L;
  memset(&L, 0, sizeof(L));
  L.dim = dim;
  L.N = 1;
  for (int d = 0; d < dim; ++d) L.nn[d] = nn[d], L.h[d] = h[d], L.N *= nn[d];
  L.m = (double*)m;
  L.g = (double*)gamma;
  assemble(&L, wre, wim, neumann, sommerfeld, bc);
  spmv(&L.A, (const cplx*)x, (cplx*)y);
  free(L.A.rp), free(L.A.ci), free(L.A.v), free(L.dinv);
  return 0;
}

/* Solve H x_c = b_c for nrhs columns one after another (the plain method).
   b, x: interleaved complex, N x nrhs column-major.  iters[c] receives the
   FGMRES iteration count; times in seconds. */
int slmg_cpu_solve(int dim, const int64_t* nn, const double* h, const double* m, const double* gamma, double wre,
                   double wim, int neumann, int sommerfeld, int bc, int levels, const double* shift, double relax,
                   int pre, int post, int cycle_w, int wlevels, int inner, int max_outer, double tol,
                   const double* b, double* x, int nrhs, int* iters, double* setup_s, double* solve_s) {
  if (dim < 2 || dim > 3 || levels < 1 || inner < 1 || nrhs < 1) return -1;
  hierarchy Hh;
  memset(&Hh, 0, sizeof(Hh));
  Hh.pre = pre, Hh.post = post, Hh.cycle_w = cycle_w, Hh.wlevels = wlevels, Hh.relax = relax;
  double t0 = now_s();
  int rc = build(&Hh, dim, nn, h, m, gamma, wre, wim, neumann, sommerfeld, bc, levels, shift);
  level F;  /* the true operator (no shift) for the outer iteration */
  memset(&F, 0, sizeof(F));
  if (rc == 0) {
    F.dim = dim;
    F.N = 1;
    for (int d = 0; d < dim; ++d) F.nn[d] = nn[d], F.h[d] = h[d], F.N *= nn[d];
    F.m = (double*)m;
    F.g = (double*)gamma;
    assemble(&F, wre, wim, neumann, sommerfeld, bc);
  }
  double t1 = now_s();
  if (rc == 0)
    for (int c = 0; c < nrhs; ++c)
      fgmres(&Hh, &F, (const cplx*)b + (int64_t)c * F.N, (cplx*)x + (int64_t)c * F.N, inner, max_outer, tol,
             &iters[c]);
  double t2 = now_s();
  if (setup_s) *setup_s = t1 - t0;
  if (solve_s) *solve_s = t2 - t1;
  free(F.A.rp), free(F.A.ci), free(F.A.v), free(F.dinv);
  free_hierarchy(&Hh);
  return rc;
}
