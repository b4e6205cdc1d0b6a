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
preconditioner, inputs resident in HBM, stencil launch shapes autotuned) with
the learned U-Net coarse-grid correction at the 257x129 level, chosen against
the plain cycle by the true residual of the first two restart cycles; the
network is trained first on pairs from plain SL-MG solves on seeded 513x513
media (training time reported, not timed).  The same batched solve with the
plain cycle, and the plain method (complex128 preconditioner, one source per
call) on a subset of sources, are reported beside it; the line's ``value`` is
the learned+batched ms/RHS.  ``roofline`` times every launch of the dominant
kernel inside the timed solves with HIP events; ``cpu_baseline`` solves one
source with the oracle's C/OpenMP restatement of the plain method;
``three_d`` times the 3-D case (193 x 193 x 97, 32 sources) with a complex64
and a complex128 preconditioner (``--no-3d`` skips it); ``stream`` solves an
uneven stream of 200 independent problems (``--no-stream`` skips it).

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
    cycle (a 4-level V cycle does not converge at 8 ppw; 4-level W saves about
    an eighth of the iterations but adds eight dense 8125-node coarsest
    solves per cycle), complex128 FGMRES(8) to 1e-6; with
    ``c64_basis`` also the complex64 preconditioner with the FGMRES basis
    stored in complex64 (``c64_c64basis_*``)."""
    import helmholtz_amd as ha
    from helmholtz_amd.device import DeviceArray
    hp, B = c3_problem()
    Bd = DeviceArray.from_numpy(B)
    out = {"workload": "3-D 193x193x97, 32 sources, 8 ppw, 5-level W(2,2), FGMRES(8), 1e-6 true residual"}
    runs = [("c64", "c128", "c64"), ("c128", "c128", "c128")]
    if c64_basis:
        runs.append(("c64", "c64", "c64_c64basis"))
    for prec, basis, key in runs:
        mg = ha.getMGparam(complex, int, 5, 1, 100, 1e-6, "Jac", 0.8, 2, 2, "W", "NoMUMPS", 0.5, 0.0)
        mg.prec = prec
        mg.basis = basis
        s = ha.getShiftedLaplacianMultigridSolver(hp, mg, 0.5, "GMRES", 8)
        t, st = timed_solves(ha, s, Bd, 1, 1, log, f"3-D {key}")
        out[f"{key}_ms_per_rhs"] = round(1e3 * t / B.shape[1], 3)
        out[f"{key}_max_iters"] = int(st["max_iters"])
        out[f"{key}_max_true_res"] = float(max(st["true_res"]))
    return out


def stream_line(log, count=200, workers=8, basis="c128"):
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
                      "max_outer": 200, "basis": basis})
        nodes += m.size * nrhs
    log(f"stream: {count} problems built in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    out = stream.solve_stream(probs, workers=workers)
    dt = time.perf_counter() - t0
    rhs = sum(int(p["B"].shape[1]) for p in probs)
    res = max(float(np.max(info["true_res"])) for _, info in out)
    ok = all(info["code"] == 0 for _, info in out) and res <= 1e-6
    log(f"stream: {rhs} right-hand sides in {dt:.2f} s, max true residual {res:.2e}")
    return {"workload": f"{count} 2-D problems, grids {{129..1025}} x {{129..513}}, 6-12 ppw, 1-64 sources each",
            "problems": count, "rhs": rhs, "node_rhs": int(nodes), "seconds": round(dt, 3),
            "ms_per_rhs": round(1e3 * dt / rhs, 3), "max_true_res": res, "all_converged": bool(ok),
            "workers": workers, "basis": basis, "buffers": "host (PCIe-inclusive)"}


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
    for k in range(warmup):
        X, solver = ha.solveLinearSystem(None, Bd, solver)
        log(f"{name} warmup {k}: max iters {solver.last_stats['max_iters']}, "
            f"fell back {solver.last_stats['fell_back']}")
    if before_timed is not None:
        before_timed(solver)
    times = []
    for k in range(steps):
        t0 = time.perf_counter()
        X, solver = ha.solveLinearSystem(None, Bd, solver)
        times.append(time.perf_counter() - t0)
        log(f"{name} step {k}: {times[-1]:.3f} s, max iters {solver.last_stats['max_iters']}, "
            f"learned used {solver.last_stats['learned_used']}, fell back {solver.last_stats['fell_back']}")
    return float(np.mean(times)), solver.last_stats


# The dominant kernel of the batched solve (rocprofv3 kernel trace of this
# benchmark): the fused up leg of the multigrid cycle on the fine grid.
ROOFLINE_KERNEL = ("up", 0)
ROOFLINE_NAME = "k_up_fused<float> on the 1025x513 grid (x += P e_c, two damped-Jacobi sweeps)"


def roofline_from_solver(ha, solver, nsrc: int, N: int, Nc: int):
    """HIP-event timing of every launch of the roofline kernel inside the timed
    solves (on the solver's stream), against its algorithmic bytes: per active
    RHS read x, b (8 B each), the coarse correction (8 B per coarse node),
    write x (8 B); plus the shared complex64 diagonal and Jacobi factor (8 B
    each per node)."""
    t = ha.kernel_time(solver)
    ms, n, byt = t["ms"], t["launches"], t["bytes"]
    achieved = byt / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    full = nsrc * (3 * N * 8 + Nc * 8) + 2 * N * 8
    traffic = _pmc_traffic()
    return {"bound": "hbm", "kernel": ROOFLINE_NAME, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "launches": n, "avg_launch_ms": round(ms / max(n, 1), 5),
            "alg_bytes_per_launch": round(byt / max(n, 1)), "alg_bytes_full_batch_launch": full}


def _pmc_traffic():
    """HBM bytes per full-batch launch of the roofline kernel from the two
    rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) summarised in
    profiles/pmc_up_fused_c2_1025x513_256rhs.md, each scaled by the counter's
    reading for a device copy of known size; None when that summary is absent."""
    f = os.path.join(ROOT, "profiles", "pmc_up_fused_c2_1025x513_256rhs.md")
    if not os.path.exists(f):
        return None
    for line in open(f):
        if line.startswith("traffic_bytes_per_launch ="):
            return float(line.split("=", 1)[1])
    return None


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
    ap.add_argument("--no-stream", action="store_true", help="skip the uneven-stream line")
    ap.add_argument("--train-media", type=int, default=40)
    ap.add_argument("--train-epochs", type=int, default=6)
    ap.add_argument("--no-c64-basis", action="store_true",
                    help="skip the solves with the FGMRES basis stored in complex64")
    args = ap.parse_args()
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

    # ---- batched, complex64 preconditioner, plain W cycle (autotuned launches)
    sb = make_solver(hp, "c64", fast=True)
    t_b, st_b = timed_solves(ha, sb, Bd, args.warmup, args.steps, log, "batched")
    res_b = np.asarray(st_b["true_res"])
    batched_ms_rhs = 1e3 * t_b / args.nsrc
    # the same solve from host buffers: B and X cross PCIe inside the call
    t0 = time.perf_counter()
    _, sb = ha.solveLinearSystem(None, B, sb)
    host_ms_rhs = 1e3 * (time.perf_counter() - t0) / args.nsrc
    log(f"batched from host buffers: {host_ms_rhs:.3f} ms/RHS")

    # ---- the same batched solve with the FGMRES basis stored in complex64
    c64b = None
    if not args.no_c64_basis:
        sc = make_solver(hp, "c64", fast=True, basis="c64")
        t_c, st_c = timed_solves(ha, sc, Bd, args.warmup, args.steps, log, "batched c64 basis")
        c64b = {"ms_per_rhs": round(1e3 * t_c / args.nsrc, 4), "max_iters": int(st_c["max_iters"]),
                "max_true_res": float(max(st_c["true_res"])),
                "note": "FGMRES basis stored in complex64; products, sums, iterate and restart residual "
                        "float64/complex128; plain cycle"}
        del sc

    # ---- learned + batched: U-Net coarse-grid correction at the 257x129 level
    from helmholtz_amd import learned, train
    net, tinfo = train.train_coarse_correction(n=(513, 513), media_count=args.train_media, level=2,
                                               epochs=args.train_epochs, log=log)
    log(f"training: {tinfo}")
    sl = make_solver(hp, "c64", fast=True)
    learned.attach(sl, net, 2)
    t_l, st_l = timed_solves(ha, sl, Bd, args.warmup, args.steps, log, "learned",
                             before_timed=lambda s: ha.time_kernel(s, *ROOFLINE_KERNEL))
    lv = ha.hierarchy_levels(sl)
    roof = roofline_from_solver(ha, sl, args.nsrc, int(np.prod(lv[0])), int(np.prod(lv[1])))
    log(f"roofline: {roof}")
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

    c3 = None if args.no_3d else c3_lines(log, not args.no_c64_basis)
    st5 = None if args.no_stream else stream_line(log)
    copy_gbs = calibration_copy(Bd, log)
    roof["copy_gbs"] = round(copy_gbs, 1)
    roof["copy_frac"] = round(roof["achieved"] / copy_gbs, 4)
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
                   "learned": "U-Net at level 2 (257x129) vs plain cycle, chosen by a two-restart trial on the first solve", "graphs": False,
                   "autotune": True, "parallelism": "rhs-batched, 1 GPU"},
        "learned_ms_per_rhs": round(learned_ms_rhs, 4),
        "learned_max_iters": int(st_l["max_iters"]),
        "learned_used": bool(st_l["learned_used"]),
        "learned_fell_back": bool(st_l["fell_back"]),
        "learned_max_true_res": float(res_l.max()),
        "training": tinfo,
        "batched_ms_per_rhs": round(batched_ms_rhs, 4),
        "batched_max_iters": int(st_b["max_iters"]),
        "batched_max_true_res": float(res_b.max()),
        "pcie_inclusive_ms_per_rhs": round(host_ms_rhs, 4),
        "c64_basis": c64b,
        "all_converged": bool(np.all(res_b <= 1e-6)) and bool(np.all(res_l <= 1e-6)) and bool(max(plain_res) <= 1e-6),
        "plain_ms_per_rhs": round(plain_ms_rhs, 3),
        "plain_iters": plain_it,
        "plain_max_true_res": float(max(plain_res)),
        "plain_rhs_sampled": args.plain_rhs,
        "speedup_vs_plain": round(plain_ms_rhs / learned_ms_rhs, 2),
        "speedup_batched_vs_plain": round(plain_ms_rhs / batched_ms_rhs, 2),
        "roofline": roof,
        "cpu_baseline": cpu,
        "three_d": c3,
        "stream": st5,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
