"""Operator side of the reference API, backed by libhelmholtz_amd on the GPU.

Mirrors Helmholtz.jl's exports for this path: ``HelmholtzParam``,
``getShiftedHelmholtzParam`` (src/Helmholtz.jl:13-34), ``GetHelmholtzOperator``
(3 methods, src/GetHelmholtz.jl:14-50), ``GetHelmholtzShiftOP`` (:81-83),
``getABL`` (:97-220), ``getMaximalFrequency`` (:75-79),
``getAcousticPointSource``/``getTopPointSrc``/``getMidPointSrc``/``loc2cs``
(src/getPointSource.jl:63-111) and jInv's ``getRegularMesh``.

Differences from the reference, by design: the operator is never assembled.
``GetHelmholtzOperator`` returns a matrix-free :class:`HelmholtzOperator`
whose ``@`` runs the HIP stencil; ``H + GetHelmholtzShiftOP(...)`` gives the
shifted operator.  Subscripts stay 1-based as in the reference; arrays are
column-major (``order='F'``), axis 0 = x, the last axis = depth.

Host vectors are numpy arrays (N,) or (N, nrhs) in Fortran order; device
blocks are :class:`helmholtz_amd.device.DeviceArray` of the same shape.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


# ------------------------------------------------------------------ mesh
@dataclass
class RegularMesh:
    """jInv.Mesh.RegularMesh: ``n`` cells per axis on ``domain`` (lo, hi pairs)."""
    domain: tuple
    n: np.ndarray
    h: np.ndarray = field(init=False)
    dim: int = field(init=False)

    def __post_init__(self):
        self.n = np.asarray(self.n, dtype=np.int64).ravel()
        dom = np.asarray(self.domain, dtype=np.float64).ravel()
        self.domain = tuple(dom)
        self.dim = len(self.n)
        self.h = np.array([(dom[2 * d + 1] - dom[2 * d]) / self.n[d] for d in range(self.dim)])

    @property
    def nodes(self) -> np.ndarray:
        return self.n + 1


def getRegularMesh(domain, n) -> RegularMesh:
    return RegularMesh(tuple(np.asarray(domain, dtype=float).ravel()), np.asarray(n))


# ------------------------------------------------------------------ parameters
@dataclass
class HelmholtzParam:
    """src/Helmholtz.jl:13-20."""
    Mesh: RegularMesh
    gamma: np.ndarray
    m: np.ndarray
    omega: complex
    NeumannOnTop: bool
    Sommerfeld: bool


def getShiftedHelmholtzParam(p: HelmholtzParam, s: float) -> HelmholtzParam:
    """src/Helmholtz.jl:32-34: gamma + s * Re(omega)."""
    return HelmholtzParam(p.Mesh, np.asarray(p.gamma, dtype=float) + s * complex(p.omega).real, p.m, p.omega,
                          p.NeumannOnTop, p.Sommerfeld)


def clear(p):
    """``clear!(HelmholtzParam)`` (src/Helmholtz.jl:25-30) is effectively a no-op
    on the record (its assignments are to locals); provided for API parity."""
    return None


# ------------------------------------------------------------------ helpers
def _i64x3(v, fill=1):
    a = (C.c_int64 * 3)(fill, fill, fill)
    for d, x in enumerate(v):
        a[d] = int(x)
    return a


def _f64x3(v, fill=1.0):
    a = (C.c_double * 3)(fill, fill, fill)
    for d, x in enumerate(v):
        a[d] = float(x)
    return a


def _is_device(x) -> bool:
    from .device import DeviceArray
    return isinstance(x, DeviceArray)


def getABL(n, NeumannAtFirstDim: bool, ABLpad, ABLamp: float, code=None):
    """src/GetHelmholtz.jl:97-220 computed on the GPU (``hh_abl``; with ``code``,
    the reference's length(n) x 2 Bool face mask, ``hh_abl_faces``: it switches
    the low/high face of each axis in 3-D and is not read in 2-D, as in the
    reference).  ``n`` are node counts; returns a float64 array shaped
    ``tuple(n)`` in Fortran order."""
    n = [int(v) for v in np.asarray(n).ravel()]
    pad = [int(v) for v in np.asarray(ABLpad).ravel()]
    out = np.zeros(int(np.prod(n)), dtype=np.float64)
    if code is None:
        L.check(L.lib().hh_abl(len(n), _i64x3(n), int(bool(NeumannAtFirstDim)), _i64x3(pad), float(ABLamp),
                               out.ctypes.data, L.HH_HOST))
    else:
        c = np.asarray(code, dtype=bool)
        if c.shape != (len(n), 2):
            raise ValueError("code must be a length(n) x 2 Bool matrix")
        faces = (C.c_int32 * 6)(*([int(v) for v in c.ravel()] + [1] * (6 - c.size)))
        L.check(L.lib().hh_abl_faces(len(n), _i64x3(n), int(bool(NeumannAtFirstDim)), _i64x3(pad), float(ABLamp),
                                     faces, out.ctypes.data, L.HH_HOST))
    return out.reshape(tuple(n), order="F")


def getSommerfeldBC(Msh: RegularMesh, mNodal, omega: float, NeumannOnTop: bool, orderNeumannBC: int = 2):
    """src/GetHelmholtz.jl:222-247 computed on the GPU (``hh_sommerfeld_bc``):
    -i omega (BC / h_d) sqrt(m) on every boundary face (the top face skipped
    with NeumannOnTop), summed at edges and corners.  ``omega`` is real, as
    the reference's ``omega::Float64``.  Returns a complex128 array shaped
    like the nodal grid, in Fortran order."""
    if isinstance(omega, complex) or np.iscomplexobj(omega):
        raise TypeError("getSommerfeldBC takes a real omega")
    nodes = [int(v) for v in Msh.nodes]
    m = np.ascontiguousarray(np.asarray(mNodal, dtype=np.float64).ravel(order="F"))
    if m.size != int(np.prod(nodes)):
        raise ValueError("mNodal must have one value per node")
    out = np.zeros(m.size, dtype=np.complex128)
    L.check(L.lib().hh_sommerfeld_bc(len(nodes), _i64x3(nodes), _f64x3(Msh.h), m.ctypes.data, float(omega),
                                     int(bool(NeumannOnTop)), int(orderNeumannBC), out.ctypes.data, L.HH_HOST))
    return out.reshape(tuple(nodes), order="F")


def getMaximalFrequency(m, M: RegularMesh) -> float:
    """src/GetHelmholtz.jl:75-79 (slowness squared in): 10 points per wavelength."""
    return (0.1 * 2 * math.pi) / (float(np.max(M.h)) * math.sqrt(float(np.max(m))))


@dataclass
class ShiftOp:
    """GetHelmholtzShiftOP(m, omega, shift) = diag(i shift omega^2 m) (src/GetHelmholtz.jl:81-83)."""
    m: np.ndarray
    omega: float
    shift: float


def GetHelmholtzShiftOP(mNodal, omega: float, shift: float) -> ShiftOp:
    return ShiftOp(np.asarray(mNodal, dtype=float), float(omega), float(shift))


# ------------------------------------------------------------------ operator
class _OpHandle:
    """Owns one ``hh_op``; shared between the shifted views of an operator."""

    def __init__(self, mesh: RegularMesh, m, omega, gamma, neumann: bool, sommerfeld: bool, order: int,
                 ho_beta=None):
        nn = [int(v) for v in mesh.nodes]
        N = int(np.prod(nn))
        self.m = np.ascontiguousarray(np.asarray(m, dtype=np.float64).ravel(order="F"))
        self.gamma = np.ascontiguousarray(np.asarray(gamma, dtype=np.float64).ravel(order="F"))
        if self.m.size != N or self.gamma.size != N:
            raise ValueError(f"m and gamma must have {N} nodal values")
        w = complex(omega)
        self.N, self.nn, self.dim = N, nn, mesh.dim
        self.omega = w
        self.h = np.array(mesh.h, dtype=np.float64)
        self.neumann, self.sommerfeld, self.order = bool(neumann), bool(sommerfeld), int(order)
        h = C.c_void_p()
        self.ho_beta = ho_beta
        if ho_beta is None:
            L.check(L.lib().hh_op_create(mesh.dim, _i64x3(nn), _f64x3(mesh.h), self.m.ctypes.data,
                                         self.gamma.ctypes.data, w.real, w.imag, int(bool(neumann)),
                                         int(bool(sommerfeld)), int(order), C.byref(h)))
        else:
            bl, bm = ho_beta
            L.check(L.lib().hh_op_create_ho(mesh.dim, _i64x3(nn), _f64x3(mesh.h), self.m.ctypes.data,
                                            self.gamma.ctypes.data, w.real, w.imag, int(bool(neumann)),
                                            int(bool(sommerfeld)), float(bl), float(bm), C.byref(h)))
        self.ptr = h

    def matches(self, p: HelmholtzParam) -> bool:
        """True if this operator was built from the fields of ``p`` (mesh,
        slowness, attenuation, frequency and both boundary flags)."""
        nn = [int(v) for v in p.Mesh.nodes]
        if self.nn != nn or not np.allclose(self.h, p.Mesh.h, rtol=1e-14, atol=0.0):
            return False
        if self.neumann != bool(p.NeumannOnTop) or self.sommerfeld != bool(p.Sommerfeld):
            return False
        if self.omega != complex(p.omega):
            return False
        m = np.asarray(p.m, dtype=np.float64).ravel(order="F")
        g = np.asarray(p.gamma, dtype=np.float64).ravel(order="F")
        return m.size == self.N and g.size == self.N and np.array_equal(m, self.m) and np.array_equal(g, self.gamma)

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p:
            try:
                L.lib().hh_op_destroy(p)
            except Exception:
                pass
            self.ptr = None


class HelmholtzOperator:
    """Matrix-free H (+ i s w^2 diag(m)).  ``H @ X`` runs ``hh_op_apply``."""

    def __init__(self, handle: _OpHandle, shift: float = 0.0, mesh=None, param=None):
        self._h = handle
        self.shift = float(shift)
        self.Mesh = mesh
        self.param = param

    @property
    def shape(self):
        return (self._h.N, self._h.N)

    def __add__(self, other):
        if isinstance(other, ShiftOp):
            if not np.allclose(np.ravel(other.m, order="F"), self._h.m):
                raise ValueError("shift operator built on a different slowness field")
            if complex(other.omega) != complex(self._h.omega.real) or self._h.omega.imag != 0.0:
                raise ValueError("GetHelmholtzShiftOP needs the operator's real omega")
            return HelmholtzOperator(self._h, self.shift + other.shift, self.Mesh, self.param)
        return NotImplemented

    __radd__ = __add__

    def apply(self, X, prec: str = "c128", out=None):
        """Y = A X.  X: numpy (N,)/(N, nrhs) (Fortran order) or a DeviceArray."""
        code = L.HH_C128 if prec == "c128" else L.HH_C64
        if _is_device(X):
            want = np.complex128 if code == L.HH_C128 else np.complex64
            if X.dtype != want or X.shape[0] != self._h.N:
                raise ValueError(f"device input must be an N x nrhs {np.dtype(want).name} block")
            nrhs = 1 if len(X.shape) == 1 else X.shape[1]
            Y = X.empty_like() if out is None else out
            L.check(L.lib().hh_op_apply(self._h.ptr, self.shift, X.ptr, Y.ptr, nrhs, code, L.HH_DEVICE))
            return Y
        dt = np.complex128 if code == L.HH_C128 else np.complex64
        Xa = np.asarray(X)
        vec = Xa.ndim == 1
        X2 = np.asfortranarray(Xa.reshape(self._h.N, -1, order="F").astype(dt, copy=False))
        Y = np.empty_like(X2, order="F")
        L.check(L.lib().hh_op_apply(self._h.ptr, self.shift, X2.ctypes.data, Y.ctypes.data, X2.shape[1], code,
                                    L.HH_HOST))
        return Y[:, 0] if vec else Y

    def __matmul__(self, X):
        return self.apply(X)

    def __mul__(self, X):
        return self.apply(X)


def getHelmholtzFun(ShiftedHT, ShiftMat, y, numCores: int = 1):
    """src/GetHelmholtz.jl:85-95: ``Hfun(x)`` writes (ShiftedHT + ShiftMat) x
    into ``y`` and returns ``y`` (the reference's two SpMatMul calls).  Both
    operands are matrix-free here: ``ShiftedHT`` a HelmholtzOperator (the
    shifted operator, e.g. ``H + GetHelmholtzShiftOP(m, w, s)``) and
    ``ShiftMat`` a ShiftOp (``GetHelmholtzShiftOP(m, w, -s)`` takes the shift
    back out), so their sum is one operator and ``Hfun`` is one
    ``hh_op_apply``.  ``y`` is a numpy array or a DeviceArray of x's shape;
    ``numCores`` is accepted for the signature.  Sparse matrices are refused:
    there is no host SpMV path."""
    if not isinstance(ShiftedHT, HelmholtzOperator) or not isinstance(ShiftMat, ShiftOp):
        raise TypeError("getHelmholtzFun takes a HelmholtzOperator and a ShiftOp (matrix-free operators only)")
    op = ShiftedHT + ShiftMat

    def Hfun(x):
        if _is_device(y):
            return op.apply(x, out=y)
        y[...] = np.asarray(op.apply(x)).reshape(y.shape, order="F")
        return y

    return Hfun


def _build(mesh, m, omega, gamma, neumann, sommerfeld, order, param=None):
    return HelmholtzOperator(_OpHandle(mesh, m, omega, gamma, neumann, sommerfeld, order), 0.0, mesh, param)


def GetHelmholtzOperator(*args, orderNeumannBC: int = 2):
    """The three reference methods (src/GetHelmholtz.jl:14-16, 22-31, 33-50):

    * ``GetHelmholtzOperator(Hparam[, order])`` -> H
    * ``GetHelmholtzOperator(Msh, m, omega, gamma, NeumannAtFirstDim, Sommerfeld[, order])`` -> H
    * ``GetHelmholtzOperator(Msh, m, omega, gamma, NeumannAtFirstDim, ABLpad, ABLamp, Sommerfeld[, order])``
      -> (H, gamma + ABL)   (gamma may be empty: then gamma = ABL)
    """
    if isinstance(args[0], HelmholtzParam):
        p = args[0]
        order = args[1] if len(args) > 1 else orderNeumannBC
        return _build(p.Mesh, p.m, p.omega, p.gamma, p.NeumannOnTop, p.Sommerfeld, order, p)
    mesh, m, omega, gamma, neumann = args[:5]
    if isinstance(args[5], (bool, np.bool_)):
        order = args[6] if len(args) > 6 else orderNeumannBC
        return _build(mesh, m, omega, gamma, neumann, args[5], order)
    pad, amp, sommerfeld = args[5], args[6], args[7]
    order = args[8] if len(args) > 8 else orderNeumannBC
    abl = getABL(mesh.nodes, neumann, pad, float(amp))
    if gamma is None or np.size(gamma) == 0:
        g = abl
    else:
        g = np.asarray(gamma, dtype=float).reshape(abl.shape, order="F") + abl
    return _build(mesh, m, omega, g, neumann, sommerfeld, order), g


def GetHelmholtzOperatorHO(*args, beta=1.0):
    """``GetHelmholtzOperatorHO`` (src/GetHelmholtz.jl:18-20, 54-72): the spread
    Laplacian + averaged mass operator, matrix-free on the GPU.

    * ``GetHelmholtzOperatorHO(Hparam[, beta])``
    * ``GetHelmholtzOperatorHO(Msh, m, omega, gamma, NeumannAtFirstDim, Sommerfeld[, beta])``

    beta: scalar in 2-D; in 3-D the reference's ``[beta_lap; beta_mass]`` (a
    scalar sets both).  Passed as ``ShiftedHT`` to ``solveLinearSystem`` it
    becomes the outer operator of the solve."""
    if isinstance(args[0], HelmholtzParam):
        p = args[0]
        b = args[1] if len(args) > 1 else beta
        mesh, m, omega, gamma, neumann, sommerfeld = p.Mesh, p.m, p.omega, p.gamma, p.NeumannOnTop, p.Sommerfeld
        param = p
    else:
        mesh, m, omega, gamma, neumann, sommerfeld = args[:6]
        b = args[6] if len(args) > 6 else beta
        param = None
    bb = np.broadcast_to(np.asarray(b, dtype=float).ravel(), (2,))
    op = HelmholtzOperator(_OpHandle(mesh, m, omega, gamma, neumann, sommerfeld, 1, (bb[0], bb[1])), 0.0, mesh,
                           param)
    return op


# ------------------------------------------------------------------ sources
def getTopPointSrc(M: RegularMesh):
    """src/getPointSource.jl:63-70 (1-based)."""
    if M.dim == 3:
        return [int(M.n[0] + 1) // 2, int(M.n[1] + 1) // 2, 1]
    return [int(M.n[0] + 1) // 2, 1]


def getMidPointSrc(M: RegularMesh):
    """src/getPointSource.jl:72-79 (1-based)."""
    return [int(v + 1) // 2 for v in M.n]


def loc2cs(*args) -> int:
    """src/getPointSource.jl:82-100: 1-based column-major linear index, in the
    reference's two methods ``loc2cs(n, sub)`` and ``loc2cs(loc1, loc2, n)``."""
    if len(args) == 3:
        return loc2cs(args[2], [args[0], args[1]])
    n, sub = args
    n = [int(v) for v in n]
    if len(sub) == 2:
        return int(sub[0] + (sub[1] - 1) * n[0])
    return int(sub[0] + (sub[1] - 1) * n[0] + (sub[2] - 1) * n[0] * n[1])


def loc2cs3D(loc1: int, loc2: int, loc3: int, n) -> int:
    """src/getPointSource.jl:87-90."""
    return loc2cs(n, [loc1, loc2, loc3])


def getAcousticPointSource(M: RegularMesh, TYPE=np.complex128, src=None):
    """src/getPointSource.jl:105-111: q = e_src / ||h||^2, returns (q, src)."""
    if src is None:
        src = getTopPointSrc(M)
    nn = [int(v) for v in M.nodes]
    flat = np.zeros(int(np.prod(nn)), dtype=TYPE)
    flat[loc2cs(nn, src) - 1] = 1.0 / float(np.linalg.norm(M.h)) ** 2
    return flat.reshape(tuple(nn), order="F"), src


def point_sources(M: RegularMesh, subs, TYPE=np.complex128, scale=None):
    """Block of point-source right-hand sides, one column per 1-based subscript
    (the batched form of getAcousticPointSource); N x len(subs), Fortran order."""
    nn = [int(v) for v in M.nodes]
    N = int(np.prod(nn))
    B = np.zeros((N, len(subs)), dtype=TYPE, order="F")
    val = 1.0 / float(np.linalg.norm(M.h)) ** 2 if scale is None else scale
    for c, s in enumerate(subs):
        B[loc2cs(nn, s) - 1, c] = val
    return B
