"""Solver side of the reference API, backed by libhelmholtz_amd on the GPU.

Mirrors ``ShiftedLaplacianMultigridSolver`` (src/ShiftedLaplacianMultigridSolver.jl:4-15),
``getShiftedLaplacianMultigridSolver`` (:24-30), ``copySolver`` (:18-22),
``solveLinearSystem`` (:33-102) and ``clear!`` (:105-109), plus the subset
of Multigrid.jl's ``getMGparam`` the reference's tests use
(test/ShiftedLaplacianTest.jl:51-64).

The multigrid hierarchy is the rediscretised shifted operator on every level
(the reference's commented geometric path, :51-62), with full-weighting
restriction, bilinear/trilinear prolongation, damped Jacobi and an FP64
direct coarsest solve; the outer method is flexible GMRES(inner).  The
``ShiftedHT`` argument of ``solveLinearSystem`` selects the operator as the
reference's does (its first solve builds from it): a matrix-free operator is
used as it is, a sparse matrix is checked against ``helmParam``, and an
operator that does not belong to ``helmParam`` is refused.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .helmholtz import HelmholtzOperator, HelmholtzParam, _OpHandle, _is_device


@dataclass
class MGparam:
    """Settings of Multigrid.jl's MGparam that this path honours.

    ``prec`` ("c128" | "c64") selects the preconditioner precision (an
    extension: the reference runs the type of the matrix).  ``basis``
    ("c128" | "c64") is the storage of the FGMRES basis; "c64" needs
    prec="c64", FGMRES and the 5/7-point operator (products, the iterate and
    the restart residual stay complex128)."""
    levels: int = 3
    numCores: int = 1
    maxOuterIter: int = 50
    relativeTol: float = 1e-6
    relaxType: str = "Jac"
    relaxParam: float = 0.8
    relaxPre: int = 2
    relaxPost: int = 2
    cycleType: str = "V"
    coarseSolveType: str = "NoMUMPS"
    extra: tuple = field(default_factory=tuple)
    prec: str = "c128"
    basis: str = "c128"
    graph: bool = False
    autotune: bool = False
    fused: bool = True
    wlevels: int = 0
    verbose: bool = False
    _hierarchy_exists: bool = False


def getMGparam(*args, **kw) -> MGparam:
    """``getMGparam([T, Ti,] levels, numCores, maxIter, relativeTol, relaxType,
    relaxParam, relaxPre, relaxPost, cycleType, coarseSolveType, ...)``."""
    a = list(args)
    while a and isinstance(a[0], type):
        a.pop(0)
    names = ["levels", "numCores", "maxOuterIter", "relativeTol", "relaxType", "relaxParam", "relaxPre",
             "relaxPost", "cycleType", "coarseSolveType"]
    vals = dict(zip(names, a))
    vals.update(kw)
    mg = MGparam(**{k: v for k, v in vals.items() if k in MGparam.__dataclass_fields__})
    mg.extra = tuple(a[len(names):])
    return mg


@dataclass
class ShiftedLaplacianMultigridSolver:
    helmParam: HelmholtzParam
    MG: MGparam
    shift: np.ndarray
    Krylov: str = "BiCGSTAB"
    inner: int = 5
    doClear: int = 0
    verbose: bool = False
    setupTime: float = 0.0
    nPrec: int = 0
    solveTime: float = 0.0
    last_stats: dict = field(default_factory=dict)
    _op: object = None
    _op_key: object = None
    _handle: object = None
    _net_ref: object = None

    def __del__(self):
        _drop_handle(self)


def _drop_handle(s):
    h = getattr(s, "_handle", None)
    if h:
        try:
            L.lib().hh_solver_destroy(h)
        except Exception:
            pass
    s._handle = None


def getShiftedLaplacianMultigridSolver(helmParam: HelmholtzParam, MG: MGparam, shift, Krylov: str = "BiCGSTAB",
                                       inner: int = 5, verbose: bool = False) -> ShiftedLaplacianMultigridSolver:
    """src/ShiftedLaplacianMultigridSolver.jl:24-30 (a scalar shift is broadcast to every level)."""
    sh = np.asarray(shift, dtype=float).ravel()
    if sh.size == 1:
        sh = np.full(MG.levels, float(sh[0]))
    return ShiftedLaplacianMultigridSolver(helmParam, MG, sh, Krylov, int(inner), 0, bool(verbose))


def copySolver(s: ShiftedLaplacianMultigridSolver) -> ShiftedLaplacianMultigridSolver:
    """src/ShiftedLaplacianMultigridSolver.jl:18-22: same settings, fresh hierarchy."""
    mg = MGparam(**{k: getattr(s.MG, k) for k in MGparam.__dataclass_fields__})
    mg._hierarchy_exists = False
    return getShiftedLaplacianMultigridSolver(s.helmParam, mg, s.shift.copy(), s.Krylov, s.inner, s.verbose)


def clear(s: ShiftedLaplacianMultigridSolver):
    """``clear!`` (src/ShiftedLaplacianMultigridSolver.jl:105-109): drop the hierarchy."""
    if s._handle:
        L.check(L.lib().hh_solver_clear(s._handle))
    s.MG._hierarchy_exists = False
    s.doClear = 0


def _sweeps(mg: MGparam, l: int):
    """Damped-Jacobi sweeps (pre, post) on level ``l`` (0 = finest).  relaxPre /
    relaxPost are counts or functions of the 1-based level (finest = 1), as in
    the reference's examples (examples/PointSourceADR/runExperiments.jl:121-122,
    ``relaxPre = l->l+1``)."""
    f = lambda v: int(v(l + 1)) if callable(v) else int(v)
    return f(mg.relaxPre), f(mg.relaxPost)


def _validate(s: ShiftedLaplacianMultigridSolver):
    mg = s.MG
    if s.Krylov not in ("GMRES", "BiCGSTAB"):
        raise ValueError(f"Krylov={s.Krylov!r}: 'GMRES' (flexible GMRES) or 'BiCGSTAB'")
    if mg.relaxType not in ("Jac", "Jac-GMRES"):
        raise NotImplementedError(f"relaxType={mg.relaxType!r}: 'Jac' or 'Jac-GMRES'")
    if str(mg.cycleType) not in ("V", "W", "K"):
        raise NotImplementedError(f"cycleType={mg.cycleType!r}: 'V', 'W' and 'K' only")
    if int(mg.levels) > 16:
        raise ValueError("at most 16 levels")
    for l in range(int(mg.levels)):
        for c in _sweeps(mg, l):
            cmax = 8 if mg.relaxType == "Jac-GMRES" else 64
            if not 0 <= c <= cmax:
                raise ValueError(f"sweep count {c} on level {l + 1}: 0..{cmax}")
    if mg.prec not in ("c128", "c64"):
        raise ValueError("MG.prec must be 'c128' or 'c64'")
    if mg.basis not in ("c128", "c64"):
        raise ValueError("MG.basis must be 'c128' or 'c64'")


def _resolve_outer(s: ShiftedLaplacianMultigridSolver, ShiftedHT):
    """The operator a solve runs on, from the ``ShiftedHT`` argument.

    The reference builds its hierarchy from the matrix it is given
    (``MGsetup(ShiftedHT, ...)``, src/ShiftedLaplacianMultigridSolver.jl:65), so
    the argument is not ignored here either:

    * a matrix-free operator from ``GetHelmholtzOperator`` (H, or the shifted
      H + GetHelmholtzShiftOP(m, w, shift[1]) of the reference's tests) or
      ``GetHelmholtzOperatorHO`` is used as it is, whatever its orderNeumannBC;
      it must have been built from helmParam's fields, and a shift other than
      0 or the solver's finest-level shift is refused;
    * a sparse matrix (the reference's ``sparse(SH')``) is checked with one
      product against helmParam's shifted operator (orderNeumannBC 2, then 1);
      a matrix that matches neither is refused;
    * ``None``: helmParam's 5/7-point operator with orderNeumannBC = 2.

    Returns (key, factory) where factory() builds or returns the _OpHandle."""
    p = s.helmParam
    if isinstance(ShiftedHT, HelmholtzOperator):
        h = ShiftedHT._h
        if not h.matches(p):
            raise ValueError("ShiftedHT was built from other fields than helmParam (mesh, m, gamma, omega or "
                             "boundary flags differ)")
        sh0 = float(np.asarray(s.shift, dtype=float).ravel()[0])
        if ShiftedHT.shift != 0.0 and abs(ShiftedHT.shift - sh0) > 1e-12 * max(1.0, abs(sh0)):
            raise ValueError(f"ShiftedHT carries shift {ShiftedHT.shift}; expected 0 or the solver's shift {sh0}")
        return ("op", id(h)), (lambda: h)
    if ShiftedHT is not None and hasattr(ShiftedHT, "tocsr") and hasattr(ShiftedHT, "shape"):
        key = ("sparse", id(ShiftedHT), id(p), id(p.m), id(p.gamma))
        return key, (lambda: _match_sparse(s, ShiftedHT))
    key = ("param", id(p), id(p.m), id(p.gamma), complex(p.omega), bool(p.NeumannOnTop), bool(p.Sommerfeld))
    return key, (lambda: _OpHandle(p.Mesh, p.m, p.omega, p.gamma, p.NeumannOnTop, p.Sommerfeld, 2))


def _match_sparse(s: ShiftedLaplacianMultigridSolver, A):
    """helmParam's operator whose shifted form equals A^H (A = sparse(SH'))
    on a random vector, trying orderNeumannBC 2 then 1."""
    p = s.helmParam
    N = int(np.prod(p.Mesh.nodes))
    rng = np.random.default_rng(12345)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    want = A.tocsr().conj().T @ x
    sh0 = float(np.asarray(s.shift, dtype=float).ravel()[0])
    for order in (2, 1):
        h = _OpHandle(p.Mesh, p.m, p.omega, p.gamma, p.NeumannOnTop, p.Sommerfeld, order)
        for shift in (sh0, 0.0):
            got = HelmholtzOperator(h, shift).apply(x)
            if np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want):
                return h
    raise ValueError("ShiftedHT does not match helmParam's shifted operator (sparse(SH') with SH = H + "
                     "GetHelmholtzShiftOP(m, w, shift[1]) and orderNeumannBC 1 or 2)")


def _ensure_handle(s: ShiftedLaplacianMultigridSolver, ShiftedHT=None):
    """Bind the solver to the operator ``ShiftedHT`` resolves to
    (``_resolve_outer``) and create the library solver on it."""
    key, factory = _resolve_outer(s, ShiftedHT)
    if s._op is None or s._op_key != key:
        _drop_handle(s)
        s._op = factory()
        s._op_key = key
    if s._handle is None:
        mg = s.MG
        sw = [_sweeps(mg, l) for l in range(int(mg.levels))]
        prm = L.MGParams(levels=int(mg.levels), pre=sw[0][0], post=sw[0][1],
                         relax=float(mg.relaxParam), cycle=ord(str(mg.cycleType)), max_outer=int(mg.maxOuterIter),
                         tol=float(mg.relativeTol), inner=int(s.inner),
                         krylov=L.HH_BICGSTAB if s.Krylov == "BiCGSTAB" else L.HH_FGMRES,
                         prec=L.HH_C64 if mg.prec == "c64" else L.HH_C128, learned=0,
                         verbose=int(bool(s.verbose)), graph=int(bool(mg.graph)),
                         autotune=int(bool(mg.autotune)), fused=int(bool(mg.fused)),
                         wlevels=int(mg.wlevels),
                         relax_type=L.HH_RELAX_JAC_GMRES if mg.relaxType == "Jac-GMRES" else L.HH_RELAX_JAC,
                         basis=L.HH_C64 if mg.basis == "c64" else L.HH_C128)
        if callable(mg.relaxPre) or callable(mg.relaxPost):
            prm.per_level = 1
            for l, (npre, npost) in enumerate(sw):
                prm.pre_l[l], prm.post_l[l] = npre, npost
        shifts = np.ascontiguousarray(np.broadcast_to(np.asarray(s.shift, dtype=float), (int(mg.levels),)))
        h = C.c_void_p()
        L.check(L.lib().hh_solver_create(s._op.ptr, C.byref(prm), shifts.ctypes.data_as(C.POINTER(C.c_double)),
                                         C.byref(h)))
        s._handle = h


_STATUS = {L.HH_CONVERGED: "converged", L.HH_MAXITER: "maxiter", L.HH_NONFINITE: "nonfinite",
           L.HH_ZERO_RHS: "zero_rhs"}


def solveLinearSystem(ShiftedHT, B, param: ShiftedLaplacianMultigridSolver, doTranspose: int = 0):
    """src/ShiftedLaplacianMultigridSolver.jl:33-102 -> (X, param).

    B: numpy (N,) or (N, nrhs) complex, or a complex128 DeviceArray (then X is
    a DeviceArray too and nothing crosses PCIe).  An all-zero host B returns
    zeros without touching the GPU (src/ShiftedLaplacianMultigridSolver.jl:40-43);
    for device blocks the library reports zero columns per RHS.  doTranspose=1
    solves the adjoint system H^H X = B (:68-80)."""
    _validate(param)
    if doTranspose not in (0, 1):
        raise ValueError("doTranspose must be 0 or 1")
    if param.doClear == 1:
        clear(param)
    N = int(np.prod(param.helmParam.Mesh.nodes))
    if ShiftedHT is not None and hasattr(ShiftedHT, "shape") and tuple(ShiftedHT.shape) != (N, N):
        raise ValueError("ShiftedHT has the wrong shape for helmParam.Mesh")
    dev = _is_device(B)
    if dev:
        if B.dtype != np.complex128 or B.shape[0] != N:
            raise ValueError("device B must be an N x nrhs complex128 DeviceArray")
        nrhs = 1 if len(B.shape) == 1 else B.shape[1]
        X = B.empty_like()
        bptr, xptr, loc = B.ptr, X.ptr, L.HH_DEVICE
    else:
        Ba = np.asarray(B)
        vec = Ba.ndim == 1
        B2 = np.asfortranarray(Ba.reshape(N, -1, order="F").astype(np.complex128, copy=False))
        nrhs = B2.shape[1]
        if np.linalg.norm(B2) == 0.0:
            return np.zeros(Ba.shape, dtype=Ba.dtype), param
        X = np.empty_like(B2, order="F")
        bptr, xptr, loc = B2.ctypes.data, X.ctypes.data, L.HH_HOST
    _ensure_handle(param, ShiftedHT)
    iters = np.zeros(nrhs, dtype=np.int32)
    status = np.zeros(nrhs, dtype=np.int32)
    res = np.zeros(nrhs, dtype=np.float64)
    st = L.SolveStats(iters.ctypes.data_as(C.POINTER(C.c_int32)), status.ctypes.data_as(C.POINTER(C.c_int32)),
                      res.ctypes.data_as(C.POINTER(C.c_double)), 0, 0, 0.0, 0.0, 0, 0, 0)
    L.check(L.lib().hh_solve(param._handle, bptr, xptr, nrhs, int(doTranspose), loc, C.byref(st)))
    param.MG._hierarchy_exists = True
    param.setupTime += st.setup_time
    param.solveTime += st.solve_time
    param.nPrec += int(st.n_prec)
    param.last_stats = {"iters": iters, "status": [_STATUS[int(v)] for v in status], "true_res": res,
                        "max_iters": int(st.max_iters), "setup_time": st.setup_time, "solve_time": st.solve_time,
                        "learned_used": bool(st.learned_used), "fell_back": bool(st.fell_back)}
    if dev:
        return X, param
    return (X[:, 0] if vec else X), param


def precondition(s: ShiftedLaplacianMultigridSolver, V):
    """One multigrid preconditioner application Z = M^-1 V (inspection hook).
    V: numpy N x nrhs complex128 (Fortran order); returns Z (complex64 or
    complex128, the preconditioner's precision)."""
    _validate(s)
    _ensure_handle(s)
    N = int(np.prod(s.helmParam.Mesh.nodes))
    V2 = np.asfortranarray(np.asarray(V, dtype=np.complex128).reshape(N, -1, order="F"))
    Z = np.empty(V2.shape, dtype=np.complex64 if s.MG.prec == "c64" else np.complex128, order="F")
    L.check(L.lib().hh_solver_precondition(s._handle, V2.ctypes.data, Z.ctypes.data, V2.shape[1], L.HH_HOST))
    return Z


def hierarchy_levels(s: ShiftedLaplacianMultigridSolver):
    """Node counts of the solver's grid levels (empty before the first solve)."""
    if s._handle is None:
        return []
    n = C.c_int()
    arr = (C.c_int64 * (16 * 3))()
    L.check(L.lib().hh_solver_levels(s._handle, C.byref(n), arr, 16))
    return [tuple(arr[3 * l: 3 * l + 3]) for l in range(n.value)]


KERNELS = {"none": L.HH_KT_NONE, "down": L.HH_KT_DOWN, "mid": L.HH_KT_MID, "up": L.HH_KT_UP,
           "apply": L.HH_KT_APPLY, "dots": L.HH_KT_DOTS, "update": L.HH_KT_UPDATE}


def time_kernel(s: ShiftedLaplacianMultigridSolver, kernel: str, level: int = 0):
    """Bracket every launch of one kernel of later solves with HIP events on
    the solver's stream (instrumentation).  kernel: 'down' | 'mid' | 'up'
    (fused multigrid legs on grid level ``level``), 'apply' (H z; not with
    MG.basis = "c64", where H z is formed inside the Arnoldi passes only),
    'dots' / 'update' (Arnoldi passes 1 and 2), or 'none'.  Resets the
    totals."""
    _validate(s)
    _ensure_handle(s)
    L.check(L.lib().hh_solver_time_kernel(s._handle, KERNELS[kernel], int(level)))


def kernel_time(s: ShiftedLaplacianMultigridSolver) -> dict:
    """Totals since ``time_kernel``: event-timed milliseconds, launches and
    algorithmic bytes of the timed launches."""
    if s._handle is None:
        return {"ms": 0.0, "launches": 0, "bytes": 0.0}
    ms, n, b = C.c_double(), C.c_int64(), C.c_double()
    L.check(L.lib().hh_solver_kernel_time(s._handle, C.byref(ms), C.byref(n), C.byref(b)))
    return {"ms": ms.value, "launches": int(n.value), "bytes": b.value}
