"""CPU oracle: textbook shifted-Laplacian multigrid + FGMRES.

TEST INFRASTRUCTURE ONLY (the checker for the HIP solver; the product path
never imports it).

Helmholtz.jl delegates the multigrid cycle and the Krylov method to two
un-vendored packages, Multigrid.jl v0.8.0 (git-tree 70508e009fd1) and
KrylovMethods.jl v0.6.0 (git-tree ceb12d552f1d); neither is present, so this
module restates the published textbook algorithms they implement, following
the control flow of ``solveLinearSystem`` (src/ShiftedLaplacianMultigridSolver.jl:33-102):

* hierarchy: the *rediscretised* shifted operator on every level, built the
  way the reference's commented geometric path does it
  (src/ShiftedLaplacianMultigridSolver.jl:51-62): the fine parameter is
  shifted with ``getShiftedHelmholtzParam`` by ``shift[1]``, each coarser level
  restricts ``m`` and ``gamma`` and adds ``shift[l+1]-shift[l]``, then calls
  ``GetHelmholtzOperator`` on the coarse mesh;
* transfers: nodal bi/tri-linear prolongation P and full-weighting restriction
  R with mirror ghosts (the operator's own boundary convention), R 1 = 1;
* smoother: damped Jacobi x <- x + w D^-1 (b - A x), or ("Jac-GMRES") a few
  GMRES steps right-preconditioned by w D^-1 (``jac_gmres``);
* coarsest level: sparse direct solve (SuperLU);
* outer: right-preconditioned flexible GMRES(inner) with restarts, x0 = 0,
  stopping on ||r|| <= tol ||b|| (the Arnoldi estimate), modified Gram-Schmidt.

Parity of iteration counts with Multigrid.jl/KrylovMethods.jl is unpinned
(their sources are unavailable); the pins here are the textbook invariants in
tests/test_oracle.py and the true residual / direct-solve agreement.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import helmholtz_ref as hr


def prolong_1d(nc_nodes: int) -> sp.csr_matrix:
    nf = 2 * nc_nodes - 1
    rows, cols, vals = [], [], []
    for I in range(nc_nodes):
        rows.append(2 * I); cols.append(I); vals.append(1.0)
    for I in range(nc_nodes - 1):
        rows += [2 * I + 1, 2 * I + 1]; cols += [I, I + 1]; vals += [0.5, 0.5]
    return sp.csr_matrix((vals, (rows, cols)), shape=(nf, nc_nodes))


def restrict_1d(nc_nodes: int) -> sp.csr_matrix:
    """Full weighting with mirror ghosts: r_c[I] = (r[2I-1] + 2 r[2I] + r[2I+1]) / 4,
    r[-1] := r[1], r[nf] := r[nf-2]."""
    nf = 2 * nc_nodes - 1
    R = sp.lil_matrix((nc_nodes, nf))
    for I in range(nc_nodes):
        f = 2 * I
        R[I, f] += 0.5
        left = f - 1 if f > 0 else 1
        right = f + 1 if f < nf - 1 else nf - 2
        R[I, left] += 0.25
        R[I, right] += 0.25
    return R.tocsr()


def kron_all(mats_by_axis):
    """kron over axes with axis 0 fastest: M = M_{d-1} (x) ... (x) M_0."""
    out = mats_by_axis[0]
    for M in mats_by_axis[1:]:
        out = sp.kron(M, out)
    return out.tocsr()


def prolongation(coarse_nodes) -> sp.csr_matrix:
    return kron_all([prolong_1d(int(n)) for n in coarse_nodes])


def restriction(coarse_nodes) -> sp.csr_matrix:
    return kron_all([restrict_1d(int(n)) for n in coarse_nodes])


def coarsen_mesh(mesh: hr.RegularMesh) -> hr.RegularMesh:
    assert all(int(c) % 2 == 0 for c in mesh.n), "cells must be even to coarsen"
    return hr.RegularMesh(mesh.domain, mesh.n // 2)


class Level:
    def __init__(self, mesh, A, m, gamma):
        self.mesh, self.A, self.m, self.gamma = mesh, A, m, gamma
        self.Dinv = 1.0 / A.diagonal()
        self.P = None  # prolongation from the next coarser level
        self.R = None
        self.lu = None


def build_hierarchy(param: hr.HelmholtzParam, levels: int, shift, order_neumann_bc: int = 2):
    """Rediscretised shifted hierarchy (src/ShiftedLaplacianMultigridSolver.jl:51-62 logic)."""
    shift = np.broadcast_to(np.asarray(shift, dtype=float), (levels,))
    wr = complex(param.omega).real
    mesh = param.Mesh
    m = np.asarray(param.m, dtype=float).ravel(order="F")
    g = np.asarray(param.gamma, dtype=float).ravel(order="F") + shift[0] * wr
    hier = []
    for l in range(levels):
        p = hr.HelmholtzParam(mesh, g, m, param.omega, param.NeumannOnTop, param.Sommerfeld)
        A = hr.operator_from_param(p, order_neumann_bc)
        hier.append(Level(mesh, A, m, g))
        if l + 1 < levels:
            cmesh = coarsen_mesh(mesh)
            R = restriction(cmesh.nodes)
            P = prolongation(cmesh.nodes)
            hier[-1].R, hier[-1].P = R, P
            m = R @ m
            g = R @ g + (shift[l + 1] - shift[l]) * wr
            mesh = cmesh
    hier[-1].lu = spla.splu(hier[-1].A.tocsc())
    return hier


def jacobi(lev: Level, x, b, w, sweeps):
    for _ in range(sweeps):
        x = x + w * lev.Dinv * (b - lev.A @ x)
    return x


def jac_gmres(lev: Level, x, b, w, steps):
    """GMRES-accelerated damped-Jacobi smoothing ("Jac-GMRES", a relaxType of
    Multigrid.jl, used by test/ShiftedLaplacianTest.jl:137-142): ``steps``
    steps of GMRES on A x = b from the current x, right-preconditioned by the
    damped Jacobi matrix M = w D^-1, i.e. the iterate minimising ||b - A x||
    over x0 + M K_steps(A M, r0).  Arnoldi with modified Gram-Schmidt; the
    small least-squares problem by numpy; an exact (zero-norm) Krylov
    direction ends the iteration early."""
    if steps <= 0:
        return x
    Minv = w * lev.Dinv
    r = b - lev.A @ x
    beta = np.linalg.norm(r)
    if beta == 0.0:
        return x
    n = r.shape[0]
    V = np.zeros((n, steps + 1), dtype=complex)
    H = np.zeros((steps + 1, steps), dtype=complex)
    V[:, 0] = r / beta
    k = steps
    for i in range(steps):
        wv = lev.A @ (Minv * V[:, i])
        for q in range(i + 1):
            H[q, i] = np.vdot(V[:, q], wv)
            wv = wv - H[q, i] * V[:, q]
        H[i + 1, i] = np.linalg.norm(wv)
        if H[i + 1, i] == 0.0:
            k = i + 1
            break
        V[:, i + 1] = wv / H[i + 1, i]
    e1 = np.zeros(k + 1, dtype=complex)
    e1[0] = beta
    y = np.linalg.lstsq(H[:k + 1, :k], e1, rcond=None)[0]
    return x + Minv * (V[:, :k] @ y)


def smooth(lev: Level, x, b, w, steps, relax="Jac"):
    """The level's smoother: damped Jacobi sweeps or Jac-GMRES steps."""
    if relax == "Jac-GMRES":
        return jac_gmres(lev, x, b, w, steps)
    return jacobi(lev, x, b, w, steps)


def kcycle_correction(hier, l, rc, w, pre, post, wlevels, relax="Jac"):
    """Coarse correction of the K-cycle (Notay & Vassilevski, Numer. Linear
    Algebra Appl. 15, 2008): two steps of a minimal-residual (GCR) iteration
    on A_{l+1} e = rc, each preconditioned by one K-cycle on level l+1.
    c1 = B rc, v1 = A c1, a1 = <v1, rc>/<v1, v1>, r2 = rc - a1 v1;
    c2 = B r2, v2 = A c2 made orthogonal to v1 (c2 alongside),
    a2 = <v2, r2>/<v2, v2>; e = a1 c1 + a2 c2 (a zero <v, v> gives a zero step)."""
    A = hier[l + 1].A
    c1 = vcycle(hier, l + 1, rc, w, pre, post, "K", wlevels, relax)
    v1 = A @ c1
    n1 = np.vdot(v1, v1).real
    a1 = np.vdot(v1, rc) / n1 if n1 > 0 else 0.0
    r2 = rc - a1 * v1
    c2 = vcycle(hier, l + 1, r2, w, pre, post, "K", wlevels, relax)
    v2 = A @ c2
    g = np.vdot(v1, v2) / n1 if n1 > 0 else 0.0
    v2 = v2 - g * v1
    c2 = c2 - g * c1
    n2 = np.vdot(v2, v2).real
    a2 = np.vdot(v2, r2) / n2 if n2 > 0 else 0.0
    return a1 * c1 + a2 * c2


def sweeps_at(v, l):
    """Sweep count on level ``l`` (0 = finest): an int, a sequence indexed by
    level, or a function of the 1-based level as Multigrid.jl's relaxPre /
    relaxPost (examples/PointSourceADR/runExperiments.jl:121-122)."""
    if callable(v):
        return int(v(l + 1))
    if np.ndim(v):
        return int(v[l])
    return int(v)


def vcycle(hier, l, b, w=0.8, pre=2, post=2, cycle="V", wlevels=0, relax="Jac"):
    """One V-, W- or K-cycle for A_l x = b from x = 0 (``wlevels`` > 0: the W
    double correction only on levels < wlevels, single below; ``pre`` /
    ``post`` per level through ``sweeps_at``).

    W: the coarse-grid correction is applied twice per visit, the second time
    from the residual recomputed on this level.  With rediscretised (not
    Galerkin) coarse operators this differs from iterating the coarse cycle
    on the restricted residual; near kh ~ 2 the coarse two-grid cycle
    diverges, and only the fine-residual form stays convergent.
    K: the coarse correction is ``kcycle_correction`` (below the next-to-last
    level; there the coarse solve is exact and one step suffices)."""
    lev = hier[l]
    if l == len(hier) - 1:
        return lev.lu.solve(b)
    npre, npost = sweeps_at(pre, l), sweeps_at(post, l)
    x = np.zeros_like(b)
    x = smooth(lev, x, b, w, npre, relax)
    if cycle == "K":
        r = b - lev.A @ x
        rc = lev.R @ r
        if l + 1 == len(hier) - 1:
            xc = hier[l + 1].lu.solve(rc)
        else:
            xc = kcycle_correction(hier, l, rc, w, pre, post, wlevels, relax)
        x = x + lev.P @ xc
        return smooth(lev, x, b, w, npost, relax)
    for _ in range(2 if (cycle == "W" and (wlevels <= 0 or l < wlevels)) else 1):
        r = b - lev.A @ x
        xc = vcycle(hier, l + 1, lev.R @ r, w, pre, post, cycle, wlevels, relax)
        x = x + lev.P @ xc
    x = smooth(lev, x, b, w, npost, relax)
    return x


def fgmres(Afun, b, M, inner=10, max_outer=50, tol=1e-6, x0=None):
    """Right-preconditioned flexible GMRES(inner) with restarts and modified
    Gram-Schmidt.  Returns (x, total_iterations, relative-residual history)."""
    n = b.shape[0]
    x = np.zeros(n, dtype=complex) if x0 is None else x0.astype(complex).copy()
    bnorm = np.linalg.norm(b)
    if bnorm == 0.0:
        return np.zeros(n, dtype=complex), 0, [0.0]
    hist = []
    total = 0
    for _ in range(max_outer):
        r = b - Afun(x)
        beta = np.linalg.norm(r)
        hist.append(beta / bnorm)
        if beta <= tol * bnorm:
            break
        V = np.zeros((n, inner + 1), dtype=complex)
        Z = np.zeros((n, inner), dtype=complex)
        H = np.zeros((inner + 1, inner), dtype=complex)
        cs = np.zeros(inner, dtype=complex)
        sn = np.zeros(inner, dtype=complex)
        gvec = np.zeros(inner + 1, dtype=complex)
        gvec[0] = beta
        V[:, 0] = r / beta
        k = 0
        for j in range(inner):
            Z[:, j] = M(V[:, j])
            w = Afun(Z[:, j])
            for i in range(j + 1):
                H[i, j] = np.vdot(V[:, i], w)
                w = w - H[i, j] * V[:, i]
            H[j + 1, j] = np.linalg.norm(w)
            if H[j + 1, j] != 0:
                V[:, j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -np.conj(sn[i]) * H[i, j] + np.conj(cs[i]) * H[i + 1, j]
                H[i, j] = t
            # rotation G = [[c, s], [-conj(s), conj(c)]] with G (a, b)^T = (|(a, b)|, 0)
            a, bb = H[j, j], H[j + 1, j]
            den = np.sqrt(abs(a) ** 2 + abs(bb) ** 2)
            cs[j] = np.conj(a) / den if den != 0 else 1.0
            sn[j] = np.conj(bb) / den if den != 0 else 0.0
            H[j, j] = den
            H[j + 1, j] = 0.0
            gvec[j + 1] = -np.conj(sn[j]) * gvec[j]
            gvec[j] = cs[j] * gvec[j]
            k = j + 1
            total += 1
            hist.append(abs(gvec[j + 1]) / bnorm)
            if abs(gvec[j + 1]) <= tol * bnorm:
                break
        y = np.linalg.solve(np.triu(H[:k, :k]), gvec[:k])
        x = x + Z[:, :k] @ y
        if abs(gvec[k]) <= tol * bnorm:
            break
    return x, total, hist


def bicgstab(Afun, b, M, max_iter=50, tol=1e-6):
    """Right-preconditioned BiCGSTAB (van der Vorst, SIAM J. Sci. Stat.
    Comput. 13, 1992), x0 = 0, shadow residual r^ = r0, stopping on the
    recursive residual ||r|| <= tol ||b||; the outer method behind
    ``Krylov="BiCGSTAB"`` (src/ShiftedLaplacianMultigridSolver.jl:91-93).
    Returns (x, iterations, relative-residual history)."""
    n = b.shape[0]
    x = np.zeros(n, dtype=complex)
    bn = np.linalg.norm(b)
    if bn == 0.0:
        return x, 0, [0.0]
    r = b.astype(complex).copy()
    rh = r.copy()
    p = np.zeros(n, dtype=complex)
    v = np.zeros(n, dtype=complex)
    rho = alpha = omega = 1.0 + 0j
    hist = [1.0]
    for it in range(max_iter):
        rho_new = np.vdot(rh, r)
        if rho_new == 0:
            return x, it, hist
        if it == 0:
            p = r.copy()
        else:
            p = r + (rho_new / rho) * (alpha / omega) * (p - omega * v)
        rho = rho_new
        ph = M(p)
        v = Afun(ph)
        d = np.vdot(rh, v)
        if d == 0:
            return x, it, hist
        alpha = rho / d
        s = r - alpha * v
        sh = M(s)
        t = Afun(sh)
        tt = np.vdot(t, t).real
        omega = np.vdot(t, s) / tt if tt > 0 else 0.0
        x = x + alpha * ph + omega * sh
        r = s - omega * t
        hist.append(np.linalg.norm(r) / bn)
        if np.linalg.norm(r) <= tol * bn:
            return x, it + 1, hist
    return x, max_iter, hist


def trapezoid_weights(mesh: hr.RegularMesh, order_neumann_bc: int = 2) -> np.ndarray:
    """Nodal weights W (1/2 per boundary axis) with W L symmetric for the
    mirror-ghost Laplacian (BC = 2); all ones for BC = 1, whose L is symmetric."""
    w = np.ones(1)
    for n in mesh.nodes:  # axis 0 fastest: kron(w_{d}, ..., w_0)
        wd = np.ones(int(n))
        if order_neumann_bc == 2:
            wd[0] = wd[-1] = 0.5
        w = np.kron(wd, w)
    return w


def solve_linear_system(param: hr.HelmholtzParam, B, levels=3, shift=0.5, inner=10, max_outer=50,
                        tol=1e-6, relax=0.8, pre=2, post=2, cycle="V", order_neumann_bc=2):
    """Plain SL-MG FGMRES solve of H X = B (column by column), returning (X, iterations)."""
    hier = build_hierarchy(param, levels, shift, order_neumann_bc)
    H = hr.operator_from_param(param, order_neumann_bc)
    B2 = np.asarray(B, dtype=complex)
    vec = B2.ndim == 1
    B2 = B2.reshape(B2.shape[0], -1)
    X = np.zeros_like(B2)
    its = []
    for c in range(B2.shape[1]):
        M = lambda v: vcycle(hier, 0, v, relax, pre, post, cycle)
        x, it, _ = fgmres(lambda v: H @ v, B2[:, c], M, inner, max_outer, tol)
        X[:, c] = x
        its.append(it)
    return (X[:, 0] if vec else X), its
