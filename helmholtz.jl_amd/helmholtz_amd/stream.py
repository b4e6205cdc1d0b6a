"""Uneven stream of independent Helmholtz problems (hh_stream_solve).

Each problem has its own grid, medium, frequency and number of sources.  The
library solves them with a pool of native worker threads, largest problem
first, one operator/solver/HIP stream per problem, so launches of different
problems overlap on the GPU.  No counterpart in Helmholtz.jl."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class Problem(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("nnodes", C.c_int64 * 3), ("h", C.c_double * 3),
        ("m", C.c_void_p), ("gamma", C.c_void_p), ("omega_re", C.c_double), ("omega_im", C.c_double),
        ("neumann_top", C.c_int32), ("sommerfeld", C.c_int32), ("bc_order", C.c_int32),
        ("params", L.MGParams), ("shift_per_level", C.c_void_p), ("B", C.c_void_p), ("X", C.c_void_p),
        ("nrhs", C.c_int64), ("loc", C.c_int32), ("iters", C.c_void_p), ("status", C.c_void_p),
        ("true_res", C.c_void_p), ("net", C.c_void_p), ("learn_level", C.c_int32),
    ]


class ProblemResult(C.Structure):
    _fields_ = [("code", C.c_int32), ("max_iters", C.c_int32), ("n_converged", C.c_int32),
                ("fell_back", C.c_int32), ("seconds", C.c_double)]


def solve_stream(problems, workers: int = 4):
    """problems: list of dicts with keys ``param`` (HelmholtzParam), ``B``
    (N x nrhs complex), and optional ``levels`` (default: coarsen to <= 561
    nodes), ``cycle`` ('W'), ``prec`` ('c64'), ``shift`` (0.5), ``inner`` (10),
    ``tol`` (1e-6), ``max_outer`` (100), ``basis`` ('c128'), ``graph``, ``autotune``, ``net``
    (a learned.UNet) and ``learn_level`` (2).  Returns [(X, info)]."""
    from .train import levels_for
    lib = L.lib()
    lib.hh_stream_solve.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p]
    lib.hh_stream_solve.restype = C.c_int
    n = len(problems)
    arr = (Problem * n)()
    res = (ProblemResult * n)()
    keep, outs = [], []
    for k, pr in enumerate(problems):
        p = pr["param"]
        nodes = [int(v) for v in p.Mesh.nodes]
        m = np.ascontiguousarray(np.asarray(p.m, dtype=np.float64).ravel(order="F"))
        g = np.ascontiguousarray(np.asarray(p.gamma, dtype=np.float64).ravel(order="F"))
        B = np.asfortranarray(np.asarray(pr["B"], dtype=np.complex128).reshape(m.size, -1, order="F"))
        X = np.empty_like(B, order="F")
        nrhs = B.shape[1]
        levels = int(pr.get("levels", levels_for(nodes) if len(nodes) == 2 else 3))
        shifts = np.full(levels, float(pr.get("shift", 0.5)))
        iters = np.zeros(nrhs, dtype=np.int32)
        status = np.zeros(nrhs, dtype=np.int32)
        tres = np.zeros(nrhs, dtype=np.float64)
        keep += [m, g, B, X, shifts, iters, status, tres]
        a = arr[k]
        a.dim = len(nodes)
        for d in range(3):
            a.nnodes[d] = nodes[d] if d < len(nodes) else 1
            a.h[d] = float(p.Mesh.h[d]) if d < len(nodes) else 1.0
        a.m, a.gamma = m.ctypes.data, g.ctypes.data
        w = complex(p.omega)
        a.omega_re, a.omega_im = w.real, w.imag
        a.neumann_top, a.sommerfeld, a.bc_order = int(bool(p.NeumannOnTop)), int(bool(p.Sommerfeld)), 2
        prec = pr.get("prec", "c64")
        a.params = L.MGParams(levels=levels, pre=2, post=2, relax=0.8, cycle=ord(pr.get("cycle", "W")),
                              max_outer=int(pr.get("max_outer", 100)), tol=float(pr.get("tol", 1e-6)),
                              inner=int(pr.get("inner", 10)), krylov=L.HH_FGMRES,
                              prec=L.HH_C64 if prec == "c64" else L.HH_C128, learned=0, verbose=0,
                              graph=int(bool(pr.get("graph", False))), autotune=int(bool(pr.get("autotune", False))),
                              fused=int(bool(pr.get("fused", True))), wlevels=int(pr.get("wlevels", 0)),
                              basis=L.HH_C64 if pr.get("basis", "c128") == "c64" else L.HH_C128)
        a.shift_per_level = shifts.ctypes.data
        a.B, a.X, a.nrhs, a.loc = B.ctypes.data, X.ctypes.data, nrhs, L.HH_HOST
        a.iters, a.status, a.true_res = iters.ctypes.data, status.ctypes.data, tres.ctypes.data
        net = pr.get("net")
        a.net = net.ptr.value if net is not None else None
        a.learn_level = int(pr.get("learn_level", 2))
        outs.append((X, iters, status, tres))
    L.check(lib.hh_stream_solve(n, C.addressof(arr), int(workers), C.addressof(res)))
    out = []
    for k in range(n):
        X, iters, status, tres = outs[k]
        r = res[k]
        out.append((X, {"code": r.code, "max_iters": r.max_iters, "n_converged": r.n_converged,
                        "fell_back": bool(r.fell_back), "seconds": r.seconds, "iters": iters, "true_res": tres}))
    return out
