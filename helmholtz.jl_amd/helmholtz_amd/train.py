"""Training pipeline for the learned coarse-grid correction.

Pairs come from the plain SL-MG itself: batched complex64 FGMRES solves on
seeded random media (layered + body generator, 1500-4500 m/s) at 6-12 points
per wavelength record the right-hand sides that reach grid level `level`; the
targets are exact solves of that level's shifted operator.  The network is
trained on the GPU by the library's own forward/backward/Adam kernels.
Weights are saved as .npz with a version tag and a SHA-256 checksum.
"""
from __future__ import annotations

import hashlib
import time

import numpy as np

from . import helmholtz as hz
from . import learned, media
from .solver import getMGparam, getShiftedLaplacianMultigridSolver, solveLinearSystem

WEIGHTS_VERSION = 1


def _medium_problem(n, seed, ppw, h=10.0, pad=20):
    rng = np.random.default_rng(seed)
    v = media.layered_body_velocity_2d(n, seed=seed, n_interfaces=int(rng.integers(6, 14)))
    m = (1.0 / v**2).ravel(order="F")
    M = hz.getRegularMesh([0.0, h * (n[0] - 1), 0.0, h * (n[1] - 1)], [n[0] - 1, n[1] - 1])
    w = media.omega_for_ppw(1500.0, h, ppw)
    _, gamma = hz.GetHelmholtzOperator(M, m, w, 0.01 * w * np.ones(m.size), True, [pad, pad], w, True)
    return M, hz.HelmholtzParam(M, gamma.ravel(order="F"), m, w, True, True)


def levels_for(n, coarsest=33 * 17):
    """Number of halvings until the grid has at most `coarsest` nodes."""
    L, a, b = 1, n[0], n[1]
    while a * b > coarsest and (a - 1) % 2 == 0 and (b - 1) % 2 == 0:
        a, b = (a - 1) // 2 + 1, (b - 1) // 2 + 1
        L += 1
    return L


def generate_pairs(n=(513, 513), media_count=8, level=2, sources=32, iters=(0, 16), seed=1000,
                   capacity=None, log=None, workers=4):
    """Capture solves of ``media_count`` seeded media run ``workers`` at a time
    (each medium's operator has its own stream, and at 513 x 513 x 32 sources
    the solve's launches are short enough that one medium leaves the GPU mostly
    idle); the targets are added to the dataset one medium after another in
    medium order, so the dataset does not depend on ``workers``."""
    from concurrent.futures import ThreadPoolExecutor
    nl = [(v - 1) // 2**level + 1 for v in n]
    cap = capacity or media_count * sources * (iters[1] - iters[0])
    data = learned.Dataset(nl[0], nl[1], cap)
    rng = np.random.default_rng(seed)
    L = levels_for(n)
    # every random draw in medium order first
    draws = []
    for k in range(media_count):
        ppw = float(rng.uniform(6.0, 12.0))
        n1 = int(n[0])
        xs = rng.integers(n1 // 10, n1 - n1 // 10, size=sources)
        zs = rng.integers(2, max(3, int(n[1]) // 3), size=sources)
        draws.append((ppw, xs, zs))

    def capture(k):
        ppw, xs, zs = draws[k]
        M, hp = _medium_problem(n, seed + k, ppw)
        mg = getMGparam(complex, int, L, 1, 30, 1e-6, "Jac", 0.8, 2, 2, "W", "NoMUMPS", 0.5, 0.0)
        mg.prec = "c64"
        s = getShiftedLaplacianMultigridSolver(hp, mg, 0.5, "GMRES", 10)
        B = hz.point_sources(M, [(int(a), int(b)) for a, b in zip(xs, zs)])
        learned.capture_begin(s, level, sources * (iters[1] - iters[0]), iters[0], iters[1])
        solveLinearSystem(None, B, s)
        return s, ppw

    workers = max(1, workers)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for g in range(0, media_count, workers):  # at most `workers` solvers alive
            done = list(ex.map(capture, range(g, min(g + workers, media_count))))
            for k, (s, ppw) in enumerate(done, start=g):
                data.add_from_solver(s, sweeps=30)
                if log:
                    log(f"medium {k}: {ppw:.1f} ppw, dataset {data.size()} pairs")
            del done
    return data


def train_coarse_correction(n=(513, 513), media_count=8, level=2, epochs=8, batch=16, lr=2e-3, seed=0,
                            log=None):
    """Returns (net, info) — info holds the data/training times and losses."""
    t0 = time.perf_counter()
    data = generate_pairs(n, media_count, level, seed=1000 + seed, log=log)
    t1 = time.perf_counter()
    net = learned.UNet(seed=seed)
    err0 = net.eval(data)
    losses = net.train(data, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    err1 = net.eval(data)
    t2 = time.perf_counter()
    info = {"pairs": data.size(), "grid": [(v - 1) // 2**level + 1 for v in n], "level": level,
            "data_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2), "epochs": epochs,
            "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
            "rel_err_untrained": round(err0, 4), "rel_err_trained": round(err1, 4)}
    return net, info


def save_weights(net: learned.UNet, path: str, meta: dict | None = None):
    theta = net.get_params()
    digest = hashlib.sha256(theta.tobytes()).hexdigest()
    np.savez(path, theta=theta, version=np.int64(WEIGHTS_VERSION), sha256=np.str_(digest),
             meta=np.str_(repr(meta or {})))


def load_weights(path: str, seed: int = 0) -> learned.UNet:
    f = np.load(path, allow_pickle=False)
    if int(f["version"]) != WEIGHTS_VERSION:
        raise ValueError("weights file version mismatch")
    theta = f["theta"]
    if hashlib.sha256(theta.tobytes()).hexdigest() != str(f["sha256"]):
        raise ValueError("weights file checksum mismatch")
    net = learned.UNet(seed=seed)
    net.set_params(theta)
    return net
