"""ctypes wrapper of the oracle's C/OpenMP restatement (oracle/slmg_cpu.c).

TEST INFRASTRUCTURE ONLY (the CPU baseline and a check of the NumPy oracle;
the product path never imports it).  Build with oracle/build_cpu.py.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import helmholtz_ref as hr

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libslmg_cpu.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built; run oracle/build_cpu.py")
        L = C.CDLL(LIB)
        p64, pd, pi = C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_int)
        i32, dbl = C.c_int, C.c_double
        L.slmg_cpu_threads.restype = i32
        L.slmg_cpu_apply.argtypes = [i32, p64, pd, pd, pd, dbl, dbl, i32, i32, i32, pd, pd]
        L.slmg_cpu_solve.argtypes = [i32, p64, pd, pd, pd, dbl, dbl, i32, i32, i32, i32, pd, dbl, i32, i32, i32,
                                     i32, i32, i32, dbl, pd, pd, i32, pi, pd, pd]
        _lib = L
    return _lib


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def _common(param: hr.HelmholtzParam):
    mesh = param.Mesh
    nn = np.ascontiguousarray(np.asarray(mesh.nodes, dtype=np.int64))
    h = np.ascontiguousarray(np.asarray(mesh.h, dtype=np.float64))
    m = np.ascontiguousarray(np.asarray(param.m, dtype=np.float64).ravel(order="F"))
    g = np.ascontiguousarray(np.asarray(param.gamma, dtype=np.float64).ravel(order="F"))
    w = complex(param.omega)
    return mesh.dim, nn, h, m, g, w


def threads() -> int:
    return int(lib().slmg_cpu_threads())


def apply(param: hr.HelmholtzParam, x, order_neumann_bc: int = 2):
    dim, nn, h, m, g, w = _common(param)
    xx = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    y = np.empty_like(xx)
    rc = lib().slmg_cpu_apply(dim, _ptr(nn, C.c_int64), _ptr(h), _ptr(m), _ptr(g), w.real, w.imag,
                              int(param.NeumannOnTop), int(param.Sommerfeld), int(order_neumann_bc),
                              _ptr(xx.view(np.float64)), _ptr(y.view(np.float64)))
    assert rc == 0
    return y


def solve(param: hr.HelmholtzParam, B, levels=3, shift=0.5, inner=10, max_outer=50, tol=1e-6, relax=0.8,
          pre=2, post=2, cycle="V", wlevels=0, order_neumann_bc=2):
    """Plain SL-MG FGMRES(inner) of H X = B, one column at a time on all
    OpenMP threads.  Returns (X, iterations, setup seconds, solve seconds)."""
    dim, nn, h, m, g, w = _common(param)
    B2 = np.asarray(B, dtype=np.complex128)
    vec = B2.ndim == 1
    B2 = B2.reshape(B2.shape[0], -1)
    nrhs = B2.shape[1]
    Bt = np.ascontiguousarray(B2.T)  # column-major N x nrhs in memory
    Xt = np.zeros_like(Bt)
    its = np.zeros(nrhs, dtype=np.int32)
    sh = np.ascontiguousarray(np.broadcast_to(np.asarray(shift, dtype=np.float64), (levels,)))
    ts, tv = C.c_double(), C.c_double()
    rc = lib().slmg_cpu_solve(dim, _ptr(nn, C.c_int64), _ptr(h), _ptr(m), _ptr(g), w.real, w.imag,
                              int(param.NeumannOnTop), int(param.Sommerfeld), int(order_neumann_bc), int(levels),
                              _ptr(sh), float(relax), int(pre), int(post), int(cycle == "W"), int(wlevels),
                              int(inner), int(max_outer), float(tol), _ptr(Bt.view(np.float64)),
                              _ptr(Xt.view(np.float64)), nrhs, _ptr(its, C.c_int), C.byref(ts), C.byref(tv))
    if rc != 0:
        raise RuntimeError(f"slmg_cpu_solve failed ({rc})")
    X = Xt.T
    return (X[:, 0] if vec else X), [int(i) for i in its], ts.value, tv.value
