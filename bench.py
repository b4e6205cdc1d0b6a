"""Benchmark: ms per right-hand side to a 1e-6 true residual on one MI355X.

Workload (the batched multi-RHS 2-D case): 1025 x 513 nodal grid, h = 10 m,
seeded layered medium (1500-4500 m/s, 12 interfaces, elliptic 4500 m/s body,
5% noise), one frequency at 8 points per wavelength at 1500 m/s, attenuation
0.01 w plus a 20-cell absorbing layer of amplitude w, Neumann top, Sommerfeld
sides, 256 point sources on depth row 2 (x = 3 + 4k).  Preconditioner: 6-level
shifted-Laplacian cycle, W(2,2) on the three finest levels and V(2,2) below,
damped Jacobi 0.8, shift 0.4 on the finest level and 0.5 below (``SHIFTS``);
outer FGMRES(10) in complex128 to ||b - H x|| <= 1e-6 ||b||.  The same batched
solve with the FGMRES basis stored in complex64 (products, sums, iterate and
restart residual still float64/complex128; ``basis = "c64"``, an option beyond
the configured complex128 outer iteration) is reported beside it as
``c64_basis``, not as ``value`` (``--no-c64-basis`` skips it).

A step = one solve of all 256 sources in one batched call (complex64
preconditioner, inputs resident in HBM, stencil launch shapes autotuned) by
the learned+batched solver: the U-Net coarse-grid correction at the 257x129
level is trained first on pairs from plain SL-MG solves on seeded 513x513
media (training time reported, not timed) and chosen against the plain cycle
by the true residual of the first solve's restart cycles 0 and 1; ``value`` is
that solver's timed ms/RHS and ``value_pipeline`` says which cycle ran (the
plain cycle when the network lost its trial).  Reported beside it:
``learned_trial`` and ``learned_every_cycle`` (the network forced into every
restart cycle, the ablation); the batched plain-cycle solve; its set-up time;
the plain method (complex128 preconditioner, one source per call) on sampled
sources and on the small 2-D workload (``c1_plain``); ``accuracy``, the error
of each of the 256 results against a complex128 1e-10 reference checked
against SciPy's sparse direct solve, with that solve's time as context;
``roofline`` (the Arnoldi pass 2 family, the largest share of an iteration)
and ``roofline_kernels`` (pass 1 and the fine-level multigrid legs), every
launch event-timed inside full solves against its algorithmic bytes;
``cpu_baseline`` (the oracle's C/OpenMP restatement of the plain method on
the host's cores, one source) with ``host`` (CPU model, core counts);
``three_d`` (193 x 193 x 97, 32 sources, complex64 vs complex128
preconditioner, the fused 3-D legs' event-timed roofline, and the plain
method one source per call; ``--no-3d`` skips it); ``stream`` and ``stream_learned``, an uneven stream of 200 independent
problems without and with the trained network (``--no-stream`` skips them).

Usage: python bench.py [--gpus 1] [--steps 2] [--warmup 1] [--plain-rhs 4]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "helmholtz.jl_amd"))

HBM_PEAK_GBS = 8000.0


def c2_problem(nsrc: int):
    import helmholtz_amd as ha
    from helmholtz_amd import media
    n = (1025, 513)
    h = 10.0
    v = media.layered_body_velocity_2d(n, seed=2)
    m = (1.0 / v**2).ravel(order="F")
    M = ha.getRegularMesh([0.0, h * (n[0] - 1), 0.0, h * (n[1] - 1)], [n[0] - 1, n[1] - 1])
    w = media.omega_for_ppw(1500.0, h, 8)
    H, gamma = ha.GetHelmholtzOperator(M, m, w, 0.01 * w * np.ones(m.size), True, [20, 20], w, True)
    hp = ha.HelmholtzParam(M, gamma.ravel(order="F"), m, w, True, True)
    B = ha.point_sources(M, media.surface_sources_2d(n, nsrc))
    return M, hp, H, B, w


# shift per level (finest first): 0.4 on the finest level, 0.5 below.  On this
# workload a uniform 0.5 needs 261 iterations, this 245; a uniform 0.4 or a
# 0.45 on the third level does not converge in 2000.
SHIFTS = (0.4, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5)


def c3_problem():
    """The 3-D case: 193 x 193 x 97 nodes, h = 10 m, the layered-plus-body
    medium extruded with seeded lateral shifts, 8 ppw at 1500 m/s, a 16-cell
    absorbing layer, 32 sources on depth index 2 in a 4 x 8 lattice."""
    import helmholtz_amd as ha
    from helmholtz_amd import media
    n, h = (193, 193, 97), 10.0
    v = media.layered_body_velocity_3d(n, seed=3)
    m = (1.0 / v**2).ravel(order="F")
    M = ha.getRegularMesh([0, h * (n[0] - 1), 0, h * (n[1] - 1), 0, h * (n[2] - 1)], [n[0] - 1, n[1] - 1, n[2] - 1])
    w = media.omega_for_ppw(1500.0, h, 8)
    _, gamma = ha.GetHelmholtzOperator(M, m, w, 0.01 * w * np.ones(m.size), True, [16, 16, 16], w, True)
    hp = ha.HelmholtzParam(M, gamma.ravel(order="F"), m, w, True, True)
    subs = [(int(round(1 + (n[0] - 1) * (i + 0.5) / 8)), int(round(1 + (n[1] - 1) * (j + 0.5) / 4)), 2)
            for j in range(4) for i in range(8)]
    return hp, ha.point_sources(M, subs)


def c3_lines(log, c64_basis=True):
    """Complex64 vs complex128 preconditioner on the 3-D case: 5-level W(2,2)
    cycle, complex128 FGMRES(8) to 1e-6; the complex64 preconditioner also
    with a 4-level W(2,2) cycle (``c64_4levelW_*``: fewer iterations, eight
    dense 8125-node coarsest solves per cycle; a 4-level V cycle does not
    converge at 8 ppw); with ``c64_basis`` also the complex64 preconditioner
    with the FGMRES basis stored in complex64 (``c64_c64basis_*``)."""
    import helmholtz_amd as ha
    from helmholtz_amd.device import DeviceArray
    hp, B = c3_problem()
    Bd = DeviceArray.from_numpy(B)
    out = {"workload": "3-D 193x193x97, 32 sources, 8 ppw, 5-level W(2,2), FGMRES(8), 1e-6 true residual"}
    runs = [("c64", "c128", "c64", 5), ("c128", "c128", "c128", 5), ("c64", "c128", "c64_4levelW", 4)]
    if c64_basis:
        runs.append(("c64", "c64", "c64_c64basis", 5))
    for prec, basis, key, levels in runs:
        mg = ha.getMGparam(complex, int, levels, 1, 100, 1e-6, "Jac", 0.8, 2, 2, "W", "NoMUMPS", 0.5, 0.0)
        mg.prec = prec
        mg.basis = basis
        s = ha.getShiftedLaplacianMultigridSolver(hp, mg, 0.5, "GMRES", 8)
        t, st, setup, _ = timed_solves(ha, s, Bd, 1, 1, log, f"3-D {key}")
        out[f"{key}_ms_per_rhs"] = round(1e3 * t / B.shape[1], 3)
        out[f"{key}_max_iters"] = int(st["max_iters"])
        out[f"{key}_max_true_res"] = float(max(st["true_res"]))
        out[f"{key}_setup_s"] = round(setup, 3)
        if key == "c64":
            # the fused 3-D legs on the finest grid, every launch of a full solve event-timed
            out["roofline_kernels"] = [kernel_roofline(ha, s, Bd, k, lev, name) for k, lev, name in ROOFLINE_KERNELS_3D]
            for r in out["roofline_kernels"]:
                log(f"3-D roofline {r['kernel'].split(' ')[0]}: {r['achieved']:.0f} GB/s ({r['frac']:.3f}), "
                    f"{r['avg_launch_ms']:.3f} ms/launch")
    # the plain method: complex128 preconditioner, one source per call
    ms, its, res, setup, _ = plain_single_rhs(ha, hp, B, [0, 13], 5, "W", 8, [0.5] * 5, log=log, name="3-D plain")
    out["plain_c128_single_rhs"] = {"ms_per_rhs": round(ms, 2), "iters": its, "max_true_res": max(res),
                                    "setup_s": round(setup, 3), "sources": [0, 13]}
    return out


def stream_line(log, count=200, workers=8, basis="c128", graph=False, net=None):
    """The uneven stream: ``count`` independent 2-D problems, grids drawn from
    {129, 257, 513, 1025} x {129, 257, 513}, 6-12 points per wavelength,
    1-64 sources each, seeded layered+body media, complex64 preconditioner,
    complex128 FGMRES(10) to 1e-6, solved by hh_stream_solve (native worker
    threads, one stream per problem, host buffers in and out)."""
    import helmholtz_amd as ha
    from helmholtz_amd import media, stream
    rng = np.random.default_rng(5)
    t0 = time.perf_counter()
    probs, nodes = [], 0
    for k in range(count):
        n = (int(rng.choice([129, 257, 513, 1025])), int(rng.choice([129, 257, 513])))
        v = media.layered_body_velocity_2d(n, seed=5000 + k)
        m = (1.0 / v**2).ravel(order="F")
        M = ha.getRegularMesh([0.0, 10.0 * (n[0] - 1), 0.0, 10.0 * (n[1] - 1)], [n[0] - 1, n[1] - 1])
        w = media.omega_for_ppw(1500.0, 10.0, float(rng.uniform(6, 12)))
        _, gam = ha.GetHelmholtzOperator(M, m, w, 0.01 * w * np.ones(m.size), True, [10, 10], w, True)
        hp = ha.HelmholtzParam(M, gam.ravel(order="F"), m, w, True, True)
        nrhs = int(rng.integers(1, 65))
        xs = rng.integers(5, n[0] - 5, size=nrhs)
        probs.append({"param": hp, "B": ha.point_sources(M, [(int(x), 2) for x in xs]), "prec": "c64",
                      "max_outer": 200, "basis": basis, "graph": graph, "net": net, "learn_level": 2})
        nodes += m.size * nrhs
    log(f"stream: {count} problems built in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    out = stream.solve_stream(probs, workers=workers)
    dt = time.perf_counter() - t0
    rhs = sum(int(p["B"].shape[1]) for p in probs)
    res = max(float(np.max(info["true_res"])) for _, info in out)
    ok = all(info["code"] == 0 for _, info in out) and res <= 1e-6
    fell = sum(1 for _, info in out if info["fell_back"])
    log(f"stream: {rhs} right-hand sides in {dt:.2f} s, max true residual {res:.2e}"
        + (f", network kept by {count - fell} of {count} problems" if net is not None else ""))
    return {"workload": f"{count} 2-D problems, grids {{129..1025}} x {{129..513}}, 6-12 ppw, 1-64 sources each",
            "problems": count, "rhs": rhs, "node_rhs": int(nodes), "seconds": round(dt, 3),
            "ms_per_rhs": round(1e3 * dt / rhs, 3), "max_true_res": res, "all_converged": bool(ok),
            "workers": workers, "basis": basis, "graphs": bool(graph), "buffers": "host (PCIe-inclusive)",
            "learned": None if net is None else {"level": 2, "fell_back": fell, "kept": count - fell}}


def make_solver(hp, prec: str, levels=6, cycle="W", tol=1e-6, fast=False, basis="c128"):
    import helmholtz_amd as ha
    mg = ha.getMGparam(complex, int, levels, 1, 200, tol, "Jac", 0.8, 2, 2, cycle, "NoMUMPS", 0.5, 0.0)
    mg.prec = prec
    mg.basis = basis
    # eager launches: at 256 RHS each launch runs for 0.1-7 ms, so graph replay
    # measured no faster, and the in-solve kernel timing needs eager launches
    mg.graph = False
    mg.autotune = fast
    mg.wlevels = 3  # W double correction on the three finest levels (kh < ~3.2), V below
    return ha.getShiftedLaplacianMultigridSolver(hp, mg, np.array(SHIFTS[:levels]), "GMRES", 10, False)


def timed_solves(ha, solver, Bd, warmup, steps, log, name, before_timed=None):
    """``warmup`` untimed solves (the first builds the hierarchy: its set-up
    time is returned separately), then ``steps`` timed solves.  Returns (mean
    seconds per timed solve, the last solve's stats, set-up seconds of the
    first call, the last solution block)."""
    setup = None
    X = None
    for k in range(warmup):
        X, solver = ha.solveLinearSystem(None, Bd, solver)
        if setup is None:
            setup = float(solver.last_stats["setup_time"])
        log(f"{name} warmup {k}: max iters {solver.last_stats['max_iters']}, "
            f"fell back {solver.last_stats['fell_back']}")
    if before_timed is not None:
        before_timed(solver)
    times = []
    for k in range(steps):
        t0 = time.perf_counter()
        X, solver = ha.solveLinearSystem(None, Bd, solver)
        times.append(time.perf_counter() - t0)
        if setup is None:
            setup = float(solver.last_stats["setup_time"])
        log(f"{name} step {k}: {times[-1]:.3f} s, max iters {solver.last_stats['max_iters']}, "
            f"learned used {solver.last_stats['learned_used']}, fell back {solver.last_stats['fell_back']}")
    return (float(np.mean(times)) if times else float("nan")), solver.last_stats, setup, X


# Kernels whose launches inside the timed solves are event-timed against their
# algorithmic bytes (hh_solver_time_kernel).  By share of a FGMRES iteration
# of the batched solve (rocprofv3 kernel trace) the Arnoldi passes lead: pass 2
# (w' = w - V h, the k_update_lazy_hz family) then pass 1 (the dots,
# k_dots_hz); the largest multigrid kernel is the fine-level up leg.
ROOFLINE_KERNELS = [
    ("update", 0, "k_update_lazy_hz<*> (Arnoldi pass 2: w = H z formed again, w' = w - V h, ||w'||)"),
    ("dots", 0, "k_dots_hz<*> (Arnoldi pass 1: w = H z from the complex64 z, h = V^H w, ||w||)"),
    ("up", 0, "k_up_fused<float> on the 1025x513 grid (x += P e_c, two damped-Jacobi sweeps)"),
    ("mid", 0, "k_mid_fused<float> on the 1025x513 grid (x2 recomputed, x' = x2 + P e_c, r, R r)"),
    ("down", 0, "k_down_fused<float> on the 1025x513 grid (two sweeps from zero, r, R r)"),
]


ROOFLINE_KERNELS_3D = [
    ("down", 0, "k_down_fused3<float,1024> on the 193x193x97 grid (two sweeps from zero, b_c = R (b - A x2); z-marching)"),
    ("mid", 0, "k_mid_fused3<float,1024> on the 193x193x97 grid (x += P e_c, b_c = R (b - A x); z-marching)"),
    ("up", 0, "k_up_fused3<float,1024> on the 193x193x97 grid (x += P e_c, two damped-Jacobi sweeps; z-marching)"),
]


def kernel_roofline(ha, solver, Bd, kernel, level, name):
    """Every launch of one kernel in a full batched solve, bracketed by HIP
    events on the solver's stream: Σ algorithmic bytes / Σ duration."""
    ha.time_kernel(solver, kernel, level)
    _, solver = ha.solveLinearSystem(None, Bd, solver)
    t = ha.kernel_time(solver)
    ha.time_kernel(solver, "none", 0)
    ms, n, byt = t["ms"], t["launches"], t["bytes"]
    achieved = byt / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    traffic = _pmc_traffic(name.split(" ")[0])
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "launches": n,
            "avg_launch_ms": round(ms / max(n, 1), 5), "alg_bytes_per_launch": round(byt / max(n, 1))}


def _pmc_traffic(kernel: str):
    """HBM bytes per launch of ``kernel`` from the two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE) summarised in profiles/, each scaled by the
    counter's reading for a device copy of known size; None when no summary
    gives it.  Summary lines read ``traffic_bytes_per_launch[<kernel>] = <B>``."""
    for f in sorted(os.listdir(os.path.join(ROOT, "profiles"))):
        if not f.endswith(".md"):
            continue
        for line in open(os.path.join(ROOT, "profiles", f)):
            if line.startswith(f"traffic_bytes_per_launch[{kernel}] ="):
                return float(line.split("=", 1)[1])
    return None


def host_info():
    """The host the CPU lines ran on: CPU model, logical CPUs, CPUs this process may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def plain_single_rhs(ha, hp, B, cols, levels, cycle, inner, shifts, wlevels=0, log=None, name="plain"):
    """The plain method: complex128 preconditioner, one source per call, the
    first call (hierarchy set-up) untimed.  Returns (ms/RHS, iterations, true
    residuals, set-up seconds, solutions)."""
    from helmholtz_amd.device import DeviceArray
    mg = ha.getMGparam(complex, int, levels, 1, 400, 1e-6, "Jac", 0.8, 2, 2, cycle, "NoMUMPS", 0.5, 0.0)
    mg.prec = "c128"
    mg.wlevels = wlevels
    s = ha.getShiftedLaplacianMultigridSolver(hp, mg, np.array(shifts[:levels]), "GMRES", inner, False)
    _, s = ha.solveLinearSystem(None, DeviceArray.from_numpy(B[:, [cols[0]]]), s)
    setup = float(s.last_stats["setup_time"])
    ts, its, res, xs = [], [], [], []
    for c in cols:
        bcol = DeviceArray.from_numpy(B[:, [c]])
        t0 = time.perf_counter()
        X, s = ha.solveLinearSystem(None, bcol, s)
        ts.append(time.perf_counter() - t0)
        its.append(int(s.last_stats["max_iters"]))
        res.append(float(s.last_stats["true_res"][0]))
        xs.append(X.to_numpy()[:, 0])
        if log:
            log(f"{name} rhs {c}: {ts[-1]:.3f} s, {its[-1]} iters, res {res[-1]:.2e}")
    return 1e3 * float(np.mean(ts)), its, res, setup, xs


def c1_problem():
    """The plain-method workload: 129 x 129 on the unit square, v = 2 + 0.3 G
    (G a unit-std Gaussian random field, correlation length 0.1, seed 1),
    w = getMaximalFrequency (10 points per wavelength), attenuation 0.01 w
    plus a 10-cell absorbing layer of amplitude w, one source at the centre."""
    import helmholtz_amd as ha
    from helmholtz_amd import media
    v = media.grf_velocity_2d((129, 129), seed=1)
    m = (1.0 / v**2).ravel(order="F")
    M = ha.getRegularMesh([0.0, 1.0, 0.0, 1.0], [128, 128])
    w = ha.getMaximalFrequency(m, M)
    _, gamma = ha.GetHelmholtzOperator(M, m, w, 0.01 * w * np.ones(m.size), False, [10, 10], w, True)
    hp = ha.HelmholtzParam(M, gamma.ravel(order="F"), m, w, False, True)
    return hp, ha.point_sources(M, [(65, 65)])


def accuracy(ha, hp, B, Xb, cols, plain_x, log):
    """Error of every batched result against a float64 reference: the
    complex128 preconditioner with FGMRES(10) to a 1e-10 true residual on all
    right-hand sides, itself checked against SciPy's sparse direct solve
    (splu, one thread) on the sampled columns ``cols``; the plain method's
    sampled solutions against splu too.  Also times splu as context."""
    import scipy.sparse.linalg as spla
    sys.path.insert(0, ROOT)
    from oracle import helmholtz_ref as hr
    from helmholtz_amd.device import DeviceArray
    M = hr.get_regular_mesh(hp.Mesh.domain, hp.Mesh.n)
    H = hr.get_helmholtz_operator7(M, hp.m, hp.omega, hp.gamma, True, True).tocsc()
    t0 = time.perf_counter()
    lu = spla.splu(H)
    t_fac = time.perf_counter() - t0
    t0 = time.perf_counter()
    Xs = lu.solve(np.ascontiguousarray(B[:, cols]))
    t_sol = (time.perf_counter() - t0) / len(cols)
    mg = ha.getMGparam(complex, int, 6, 1, 400, 1e-10, "Jac", 0.8, 2, 2, "W", "NoMUMPS", 0.5, 0.0)
    mg.prec = "c128"
    mg.wlevels = 3
    sr = ha.getShiftedLaplacianMultigridSolver(hp, mg, np.array(SHIFTS[:6]), "GMRES", 10, False)
    Xr, sr = ha.solveLinearSystem(None, DeviceArray.from_numpy(B), sr)
    Xr = Xr.to_numpy()
    Xb = Xb.to_numpy()
    nref = np.linalg.norm(Xr, axis=0)
    err = np.linalg.norm(Xb - Xr, axis=0) / nref
    ref_vs_splu = [float(np.linalg.norm(Xr[:, c] - Xs[:, k]) / np.linalg.norm(Xs[:, k])) for k, c in enumerate(cols)]
    plain_err = [float(np.linalg.norm(x - Xs[:, k]) / np.linalg.norm(Xs[:, k])) for k, x in enumerate(plain_x)]
    log(f"accuracy: batched vs 1e-10 reference max {err.max():.2e}; reference vs splu max {max(ref_vs_splu):.2e}")
    return {"reference": "complex128 preconditioner, FGMRES(10) to 1e-10 true residual, all 256 sources; "
                         "checked against scipy splu on the sampled sources",
            "batched_err_max": float(err.max()), "batched_err_median": float(np.median(err)),
            "batched_err_per_rhs": [float(f"{e:.3e}") for e in err],
            "reference_vs_splu": ref_vs_splu, "sampled_sources": [int(c) for c in cols],
            "plain_err_vs_splu": plain_err,
            "splu": {"factor_s": round(t_fac, 2), "solve_s_per_rhs": round(t_sol, 3), "threads": 1,
                     "ms_per_rhs_256": round(1e3 * (t_fac + B.shape[1] * t_sol) / B.shape[1], 2),
                     "note": "SciPy SuperLU direct solve of the same 1025x513 operator on the host (context)"}}


def calibration_copy(Bd, log):
    """Device-to-device copies of the RHS block (bytes known exactly): the
    scale for the HBM byte counters of the --pmc runs (the last copy), and the
    measured copy bandwidth (read + written bytes over the best of three
    synchronous copies) that the roofline kernel is also compared with."""
    from helmholtz_amd import _lib as L
    from helmholtz_amd.device import DeviceArray
    tmp = DeviceArray(Bd.shape, Bd.dtype)
    probe = np.zeros(1, dtype=np.complex128)
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        tmp.copy_from(Bd)
        # a device-to-host read of one value completes only after the copy
        L.check(L.lib().hh_memcpy(probe.ctypes.data, tmp.ptr, probe.nbytes, 1))
        best = min(best, time.perf_counter() - t0)
    gbs = 2.0 * Bd.nbytes / best / 1e9
    log(f"calibration copy: {Bd.nbytes} bytes read, {Bd.nbytes} bytes written (device to device), {gbs:.0f} GB/s")
    del tmp
    return gbs


def cpu_baseline(hp, B, iters_per_rhs: int, budget_s: float):
    """The plain method on the host: the oracle's C/OpenMP restatement
    (assembled CSR operator, complex128 W(2,2)/V SL-MG with the same hierarchy,
    FGMRES(10)) on all OpenMP threads.  One source is solved to 1e-6 when a
    10-iteration probe predicts it fits in ``budget_s``; otherwise the probe's
    time per iteration is scaled by the GPU plain method's iterations/RHS."""
    sys.path.insert(0, ROOT)
    from oracle import helmholtz_ref as hr
    if not os.path.exists(os.path.join(ROOT, "oracle", "libslmg_cpu.so")):
        from oracle import build_cpu
        build_cpu.build()
    from oracle import slmg_cpu
    M = hr.get_regular_mesh(hp.Mesh.domain, hp.Mesh.n)
    p = hr.HelmholtzParam(M, hp.gamma, hp.m, hp.omega, True, True)
    b = B[:, 1].copy()
    kw = dict(levels=6, shift=np.array(SHIFTS[:6]), inner=10, cycle="W", wlevels=3)
    _, its, _, tprobe = slmg_cpu.solve(p, b, max_outer=1, tol=1e-30, **kw)
    per_it = tprobe / max(its[0], 1)
    nthr = slmg_cpu.threads()
    if per_it * iters_per_rhs * 1.3 <= budget_s:
        X, its, tsetup, tsolve = slmg_cpu.solve(p, b, max_outer=60, tol=1e-6, **kw)
        Hc = hr.operator_from_param(p)
        res = float(np.linalg.norm(Hc @ X - b) / np.linalg.norm(b))
        return {"value": round(tsolve * 1e3, 2), "unit": "ms/RHS", "cores": nthr, "kind": "port",
                "sample": f"C/OpenMP restatement of the plain method (CSR, c128 W/V(2,2) 6 levels, FGMRES(10)), "
                          f"source 2 of 256 solved to {res:.2e} true residual in {its[0]} iterations, "
                          f"{tsolve:.1f} s, setup {tsetup:.2f} s not counted"}
    return {"value": round(per_it * iters_per_rhs * 1e3, 2), "unit": "ms/RHS", "cores": nthr, "kind": "port",
            "sample": f"C/OpenMP restatement of the plain method, source 2 of 256, {its[0]} iterations in "
                      f"{tprobe:.1f} s, scaled by {iters_per_rhs} iterations/RHS"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsrc", type=int, default=256)
    ap.add_argument("--plain-rhs", type=int, default=4)
    ap.add_argument("--cpu-budget", type=float, default=40.0,
                    help="seconds the host baseline may spend on a full single-source solve")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3-D complex64 vs complex128 lines")
    ap.add_argument("--no-stream", action="store_true", help="skip the uneven-stream lines")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the per-result error and splu lines")
    ap.add_argument("--train-media", type=int, default=40)
    ap.add_argument("--train-epochs", type=int, default=6)
    ap.add_argument("--no-c64-basis", action="store_true",
                    help="skip the solves with the FGMRES basis stored in complex64")
    ap.add_argument("--profile-batched", type=int, default=0, metavar="CYCLES",
                    help="profiling mode: only the batched complex64 solve of the 1025x513 workload, stopped "
                         "after CYCLES FGMRES restart cycles (no warmup, no other lines)")
    ap.add_argument("--profile-3d", type=int, default=0, metavar="CYCLES",
                    help="profiling mode: only the 3-D complex64 solve (193x193x97, 32 sources), stopped after "
                         "CYCLES FGMRES(8) restart cycles")
    ap.add_argument("--profile-training", action="store_true",
                    help="profiling mode: only the training of the learned correction (pairs from 513x513 media)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the solve shards by right-hand side "
              "with no communication, so N GPUs would each run this benchmark on nsrc/N sources.",
              file=sys.stderr)
        sys.exit(2)
    import helmholtz_amd as ha
    from helmholtz_amd import learned, train
    from helmholtz_amd.device import DeviceArray

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if args.profile_3d:
        hp3, B3 = c3_problem()
        mg = ha.getMGparam(complex, int, 5, 1, args.profile_3d, 1e-6, "Jac", 0.8, 2, 2, "W", "NoMUMPS", 0.5, 0.0)
        mg.prec = "c64"
        s3 = ha.getShiftedLaplacianMultigridSolver(hp3, mg, 0.5, "GMRES", 8)
        t0 = time.perf_counter()
        _, s3 = ha.solveLinearSystem(None, DeviceArray.from_numpy(B3), s3)
        log(f"profiling 3-D solve: {args.profile_3d} restart cycles, {s3.last_stats['max_iters']} iterations, "
            f"{time.perf_counter() - t0:.2f} s")
        return
    if args.profile_training:
        _, info = train.train_coarse_correction(n=(513, 513), media_count=args.train_media, level=2,
                                                epochs=args.train_epochs, log=log)
        log(f"training: {info}")
        return
    t = time.perf_counter()
    M, hp, H, B, w = c2_problem(args.nsrc)
    log(f"problem built in {time.perf_counter() - t:.1f} s: N={H.shape[0]} nsrc={args.nsrc} omega={w:.3f}")
    Bd = DeviceArray.from_numpy(B)
    if args.profile_batched:
        sp = make_solver(hp, "c64", fast=True)
        sp.MG.maxOuterIter = args.profile_batched
        t0 = time.perf_counter()
        _, sp = ha.solveLinearSystem(None, Bd, sp)
        log(f"profiling solve: {args.profile_batched} restart cycles, {sp.last_stats['max_iters']} iterations, "
            f"{time.perf_counter() - t0:.2f} s")
        calibration_copy(Bd, log)  # a device copy of known size: the scale of the byte counters
        return

    # ---- batched, complex64 preconditioner, plain cycle (autotuned launches)
    sb = make_solver(hp, "c64", fast=True)
    # warm-up solves (the first builds the hierarchy); this solver's timed
    # steps are taken below, interleaved with the configured pipeline's
    _, st_b, setup_b, Xb = timed_solves(ha, sb, Bd, max(1, args.warmup), 0, log, "batched")
    # the same solve from host buffers: B and X cross PCIe inside the call
    t0 = time.perf_counter()
    _, sb = ha.solveLinearSystem(None, B, sb)
    host_ms_rhs = 1e3 * (time.perf_counter() - t0) / args.nsrc
    log(f"batched from host buffers: {host_ms_rhs:.3f} ms/RHS")
    # every launch of the leading kernels, event-timed inside full solves
    roofs = [kernel_roofline(ha, sb, Bd, k, l, name) for k, l, name in ROOFLINE_KERNELS]
    for r in roofs:
        log(f"roofline: {r}")

    # ---- the same batched solve with the FGMRES basis stored in complex64
    c64b = None
    if not args.no_c64_basis:
        sc = make_solver(hp, "c64", fast=True, basis="c64")
        t_c, st_c, _, _ = timed_solves(ha, sc, Bd, args.warmup, args.steps, log, "batched c64 basis")
        c64b = {"ms_per_rhs": round(1e3 * t_c / args.nsrc, 4), "max_iters": int(st_c["max_iters"]),
                "max_true_res": float(max(st_c["true_res"])),
                "note": "FGMRES basis stored in complex64; products, sums, iterate and restart residual "
                        "float64/complex128; plain cycle"}
        del sc

    # ---- learned + batched: U-Net coarse-grid correction at the 257x129 level,
    # trained on pairs from plain SL-MG solves on seeded 513x513 media
    t0 = time.perf_counter()
    net, tinfo = train.train_coarse_correction(n=(513, 513), media_count=args.train_media, level=2,
                                               epochs=args.train_epochs, log=log)
    log(f"training ({time.perf_counter() - t0:.1f} s): {tinfo}")
    sl = make_solver(hp, "c64", fast=True)
    learned.attach(sl, net, 2)
    t0 = time.perf_counter()
    _, sl = ha.solveLinearSystem(None, Bd, sl)  # builds the hierarchy; restart cycles 0/1 are the trial
    trial = {"seconds": round(time.perf_counter() - t0, 3), "setup_s": round(sl.last_stats["setup_time"], 3),
             "max_iters": int(sl.last_stats["max_iters"]), "network_ran": bool(sl.last_stats["learned_used"]),
             "kept": not sl.last_stats["fell_back"]}
    log(f"learned trial: {trial}")
    # the configured pipeline's timed steps, each followed by a solve of the
    # plain batched solver, so that both lines see the same state of the box
    # (over the minutes of this script the same solve drifts by a few per cent)
    tl, tb = [], []
    for k in range(args.steps):
        t0 = time.perf_counter()
        _, sl = ha.solveLinearSystem(None, Bd, sl)
        tl.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        _, sb = ha.solveLinearSystem(None, Bd, sb)
        tb.append(time.perf_counter() - t0)
        log(f"learned+batched step {k}: {tl[-1]:.3f} s, max iters {sl.last_stats['max_iters']}, learned used "
            f"{sl.last_stats['learned_used']}, fell back {sl.last_stats['fell_back']}; batched {tb[-1]:.3f} s, "
            f"max iters {sb.last_stats['max_iters']}")
    t_l, st_l = float(np.mean(tl)), sl.last_stats
    t_b, st_b = float(np.mean(tb)), sb.last_stats
    res_b = np.asarray(st_b["true_res"])
    batched_ms_rhs = 1e3 * t_b / args.nsrc
    del sb
    learned_used = bool(st_l["learned_used"])
    # ablation: the network in every restart cycle (no trial), at most 30
    # restart cycles (set-up time subtracted)
    sa = make_solver(hp, "c64", fast=True)
    sa.MG.maxOuterIter = 30
    learned.attach(sa, net, 2)
    learned.policy(sa, always=True)
    t_a, st_a, setup_a, _ = timed_solves(ha, sa, Bd, 0, 1, log, "network every cycle")
    t_a -= setup_a
    ablation = {"seconds": round(t_a, 3), "max_iters": int(st_a["max_iters"]),
                "max_true_res": float(max(st_a["true_res"])), "network_ran": bool(st_a["learned_used"]),
                "converged": bool(max(st_a["true_res"]) <= 1e-6), "max_restart_cycles": 30}
    ablation["ms_per_rhs"] = round(1e3 * t_a / args.nsrc, 4) if ablation["converged"] else None
    del sa
    pipeline_ms_rhs = 1e3 * t_l / args.nsrc
    ms_step = 1e3 * t_l
    del sl

    # ---- plain FP64, one source per call, and the small 2-D plain workload
    step = max(1, args.nsrc // max(1, args.plain_rhs))
    cols = [min(args.nsrc - 1, c * step + step // 2) for c in range(args.plain_rhs)]
    plain_ms_rhs, plain_it, plain_res, setup_p, plain_x = plain_single_rhs(ha, hp, B, cols, 6, "W", 10, SHIFTS,
                                                                            wlevels=3, log=log)
    hp1, B1 = c1_problem()
    c1_ms, c1_it, c1_res, c1_setup, _ = plain_single_rhs(ha, hp1, B1, [0], 3, "V", 10, [0.5] * 3, log=log,
                                                           name="C1 plain")
    c1 = {"workload": "2-D 129x129 unit square, v = 2 + 0.3 GRF, 10 ppw, 3-level V(2,2), FGMRES(10)",
          "ms_per_rhs": round(c1_ms, 3), "iters": c1_it[0], "true_res": c1_res[0], "setup_s": round(c1_setup, 4)}

    acc = None if args.no_accuracy else accuracy(ha, hp, B, Xb, cols, plain_x, log)
    del Xb
    c3 = None if args.no_3d else c3_lines(log, not args.no_c64_basis)
    st5 = st5l = None
    if not args.no_stream:
        st5 = stream_line(log)
        st5l = stream_line(log, net=net)
    copy_gbs = calibration_copy(Bd, log)
    for r in roofs:
        r["copy_frac"] = round(r["achieved"] / copy_gbs, 4)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(hp, B, int(np.mean(plain_it)), args.cpu_budget)
        log(f"cpu baseline: {cpu}")
    host = host_info()

    line = {
        "metric": "ms/RHS to 1e-6 true residual: learned+batched vs plain SL-MG, same GPU",
        "value": round(pipeline_ms_rhs, 4),
        "unit": "ms/RHS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 preconditioner / c128 outer",
        "data": "synthetic (seeded layered+body medium, point sources)",
        "config": {"workload": "2-D 1025x513 Helmholtz, 256 point sources, 8 ppw, 1e-6 true residual",
                   "nsrc": args.nsrc, "levels": 6, "cycle": "W(2,2) on levels 0-2, V(2,2) below",
                   "krylov": "FGMRES(10)",
                   "learned": "U-Net at level 2 (257x129) vs plain cycle, chosen by a two-restart trial on the "
                              "first solve", "graphs": False, "autotune": True, "parallelism": "rhs-batched, 1 GPU"},
        # value is the learned+batched solver's timed solves: with the network
        # when it won its trial, else the plain cycle it fell back to
        "value_pipeline": "batched, learned correction" if learned_used
                          else "batched, plain cycle (the network lost its trial and the solver fell back)",
        "learned_used": learned_used,
        "learned_ms_per_rhs": round(pipeline_ms_rhs, 4) if learned_used else None,
        "learned_trial": trial,
        "learned_every_cycle": ablation,
        "training": tinfo,
        "batched_ms_per_rhs": round(batched_ms_rhs, 4),
        "batched_max_iters": int(st_b["max_iters"]),
        "batched_max_true_res": float(res_b.max()),
        "batched_setup_s": round(setup_b, 3),
        "pcie_inclusive_ms_per_rhs": round(host_ms_rhs, 4),
        "c64_basis": c64b,
        "all_converged": bool(np.all(res_b <= 1e-6)) and bool(max(st_l["true_res"]) <= 1e-6)
                         and bool(max(plain_res) <= 1e-6),
        "plain_ms_per_rhs": round(plain_ms_rhs, 3),
        "plain_iters": plain_it,
        "plain_max_true_res": float(max(plain_res)),
        "plain_rhs_sampled": args.plain_rhs,
        "plain_setup_s": round(setup_p, 3),
        "speedup_vs_plain": round(plain_ms_rhs / pipeline_ms_rhs, 2),
        "speedup_batched_vs_plain": round(plain_ms_rhs / batched_ms_rhs, 2),
        "accuracy": acc,
        "c1_plain": c1,
        "roofline": roofs[0],
        "roofline_kernels": roofs[1:],
        "cpu_baseline": cpu,
        "host": host,
        "three_d": c3,
        "stream": st5,
        "stream_learned": st5l,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
