"""Benchmark: ms per right-hand side to a 1e-6 true residual on one MI355X.

Workload (the batched multi-RHS 2-D case): 1025 x 513 nodal grid, h = 10 m,
seeded layered medium (1500-4500 m/s, 12 interfaces, elliptic 4500 m/s body,
5% noise), one frequency at 8 points per wavelength at 1500 m/s, attenuation
0.01 w plus a 20-cell absorbing layer of amplitude w, Neumann top, Sommerfeld
sides, 256 point sources on depth row 2 (x = 3 + 4k).  Preconditioner: 6-level
shifted-Laplacian cycle, W(2,2) on the three finest levels and V(2,2) below,
damped Jacobi 0.8, shift 0.5 on every level;
outer FGMRES(10) in complex128 to ||b - H x|| <= 1e-6 ||b||.

A step = one solve of all 256 sources in one batched call (complex64
preconditioner, inputs resident in HBM, FGMRES iterations replayed as
hipGraphs, stencil launch shapes autotuned) with the learned U-Net coarse-grid
correction at the 257x129 level and the plain cycle as its fallback; the
network is trained first on pairs from plain SL-MG solves on seeded 513x513
media (training time reported, not timed).  The same batched solve with the
plain cycle, and the plain method (complex128 preconditioner, one source per
call, eager launches) on a subset of sources, are reported beside it; the
line's ``value`` is the learned+batched ms/RHS.

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


GRAPHS = True


def make_solver(hp, prec: str, levels=6, cycle="W", tol=1e-6, fast=False):
    import helmholtz_amd as ha
    mg = ha.getMGparam(complex, int, levels, 1, 200, tol, "Jac", 0.8, 2, 2, cycle, "NoMUMPS", 0.5, 0.0)
    mg.prec = prec
    mg.graph = fast and GRAPHS
    mg.autotune = fast
    mg.wlevels = 3  # W double correction on the three finest levels (kh < ~3.2), V below
    return ha.getShiftedLaplacianMultigridSolver(hp, mg, 0.5, "GMRES", 10, False)


def timed_solves(ha, solver, Bd, warmup, steps, log, name):
    for k in range(warmup):
        X, solver = ha.solveLinearSystem(None, Bd, solver)
        log(f"{name} warmup {k}: max iters {solver.last_stats['max_iters']}, "
            f"fell back {solver.last_stats['fell_back']}")
    times = []
    for k in range(steps):
        t0 = time.perf_counter()
        X, solver = ha.solveLinearSystem(None, Bd, solver)
        times.append(time.perf_counter() - t0)
        log(f"{name} step {k}: {times[-1]:.3f} s, max iters {solver.last_stats['max_iters']}, "
            f"fell back {solver.last_stats['fell_back']}")
    return float(np.mean(times)), solver.last_stats


def roofline_jacobi(H, nrhs: int, reps: int = 20):
    """Event-timed damped-Jacobi sweep (complex64, the batched smoother) at the
    workload's fine grid with ``nrhs`` columns.  Algorithmic bytes per node:
    read x, b, write x (3 * 8 B per RHS) + the complex64 diagonal (8 B)."""
    import ctypes as C
    from helmholtz_amd import _lib as L
    from helmholtz_amd.device import DeviceArray
    N = H.shape[0]
    rng = np.random.default_rng(0)
    col = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    X = DeviceArray.from_numpy(np.asfortranarray(np.repeat(col[:, None], nrhs, axis=1)))
    Bd = DeviceArray.from_numpy(np.asfortranarray(np.repeat(col[:, None], nrhs, axis=1)))
    Y = X.empty_like()
    ms = C.c_double()
    L.check(L.lib().hh_op_time_kernel(H._h.ptr, 2, 0.5, X.ptr, Bd.ptr, Y.ptr, nrhs, L.HH_C64, reps, C.byref(ms)))
    bytes_alg = N * (3 * 8 * nrhs + 8)
    achieved = bytes_alg / (ms.value * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "k_stencil2d<float,float,float,M_JACOBI>", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_launch": bytes_alg, "avg_launch_ms": round(ms.value, 5), "nrhs": nrhs}


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
    kw = dict(levels=6, shift=0.5, inner=10, cycle="W", wlevels=3)
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
    ap.add_argument("--train-media", type=int, default=8)
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch FGMRES iterations eagerly (for profilers that do not trace hipGraph replays)")
    args = ap.parse_args()
    global GRAPHS
    GRAPHS = not args.no_graphs
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the solve shards by right-hand side "
              "with no communication, so N GPUs would each run this benchmark on nsrc/N sources.",
              file=sys.stderr)
        sys.exit(2)
    import helmholtz_amd as ha
    from helmholtz_amd.device import DeviceArray

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t = time.perf_counter()
    M, hp, H, B, w = c2_problem(args.nsrc)
    log(f"problem built in {time.perf_counter() - t:.1f} s: N={H.shape[0]} nsrc={args.nsrc} omega={w:.3f}")
    Bd = DeviceArray.from_numpy(B)

    # ---- batched, complex64 preconditioner, plain W cycle (graphs + autotuned launches)
    sb = make_solver(hp, "c64", fast=True)
    t_b, st_b = timed_solves(ha, sb, Bd, args.warmup, args.steps, log, "batched")
    res_b = np.asarray(st_b["true_res"])
    batched_ms_rhs = 1e3 * t_b / args.nsrc

    # ---- learned + batched: U-Net coarse-grid correction at the 257x129 level
    from helmholtz_amd import learned, train
    net, tinfo = train.train_coarse_correction(n=(513, 513), media_count=args.train_media, level=2,
                                               epochs=args.train_epochs, log=log)
    log(f"training: {tinfo}")
    sl = make_solver(hp, "c64", fast=True)
    learned.attach(sl, net, 2)
    t_l, st_l = timed_solves(ha, sl, Bd, args.warmup, args.steps, log, "learned")
    res_l = np.asarray(st_l["true_res"])
    learned_ms_rhs = 1e3 * t_l / args.nsrc
    ms_step = 1e3 * t_l
    st = st_l

    # ---- plain FP64, one source per call
    sp_ = make_solver(hp, "c128")
    plain_t, plain_it, plain_res = [], [], []
    step = max(1, args.nsrc // max(1, args.plain_rhs))
    for c in range(args.plain_rhs + 1):
        col = 0 if c == 0 else min(args.nsrc - 1, (c - 1) * step + step // 2)
        bcol = DeviceArray.from_numpy(B[:, [col]])
        t0 = time.perf_counter()
        Xp, sp_ = ha.solveLinearSystem(None, bcol, sp_)
        dt = time.perf_counter() - t0
        if c == 0:
            continue  # first call builds the hierarchy
        plain_t.append(dt)
        plain_it.append(int(sp_.last_stats["max_iters"]))
        plain_res.append(float(sp_.last_stats["true_res"][0]))
        log(f"plain rhs {c}: {dt:.3f} s, {plain_it[-1]} iters, res {plain_res[-1]:.2e}")
    plain_ms_rhs = 1e3 * float(np.mean(plain_t))

    roof = roofline_jacobi(H, args.nsrc)
    log(f"roofline: {roof}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(hp, B, int(np.mean(plain_it)), args.cpu_budget)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "ms/RHS to 1e-6 true residual: learned+batched vs plain SL-MG, same GPU",
        "value": round(learned_ms_rhs, 4),
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
                   "nsrc": args.nsrc, "levels": 6, "cycle": "W(2,2) on levels 0-2, V(2,2) below", "krylov": "FGMRES(10)",
                   "learned": "U-Net at level 2 (257x129), plain-cycle fallback", "graphs": GRAPHS,
                   "autotune": True, "parallelism": "rhs-batched, 1 GPU"},
        "learned_ms_per_rhs": round(learned_ms_rhs, 4),
        "learned_max_iters": int(st_l["max_iters"]),
        "learned_fell_back": bool(st_l["fell_back"]),
        "learned_max_true_res": float(res_l.max()),
        "training": tinfo,
        "batched_ms_per_rhs": round(batched_ms_rhs, 4),
        "batched_max_iters": int(st_b["max_iters"]),
        "batched_max_true_res": float(res_b.max()),
        "all_converged": bool(np.all(res_b <= 1e-6)) and bool(np.all(res_l <= 1e-6)) and bool(max(plain_res) <= 1e-6),
        "plain_ms_per_rhs": round(plain_ms_rhs, 3),
        "plain_iters": plain_it,
        "plain_max_true_res": float(max(plain_res)),
        "plain_rhs_sampled": args.plain_rhs,
        "speedup_vs_plain": round(plain_ms_rhs / learned_ms_rhs, 2),
        "speedup_batched_vs_plain": round(plain_ms_rhs / batched_ms_rhs, 2),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
