"""GPU U-Net (BF16 MFMA, FP32 accumulate) against the oracle.

Bars (BF16 activations and weights, FP32 accumulation):
* forward: ||y_gpu - y_ref|| / ||y_ref|| <= 3e-2 against the float64 oracle on
  the same BF16-rounded parameters and inputs;
* gradient: ||g_gpu - g_ref|| / ||g_ref|| <= 6e-2 against float64 autograd of
  the independent torch statement, and per-layer cosine similarity >= 0.99;
* Adam: one step equals the closed form to float32 rounding;
* training on pairs the plain SL-MG generates lowers the loss;
* inside FGMRES the learned correction reaches the same 1e-6 true residual
  (the iteration counts with and without it are printed: that is the gain).
"""
import numpy as np
import pytest

from oracle import helmholtz_ref as hr, unet_ref as ur
from unet_torch_ref import torch_loss_grad

pytestmark = pytest.mark.gpu

import helmholtz_amd as ha  # noqa: E402
from helmholtz_amd import learned, media  # noqa: E402


def _case(seed, n=2, nx=17, ny=9):
    rng = np.random.default_rng(seed)
    feats = ur.bf16_round(rng.standard_normal((n, nx * ny, 4)).astype(np.float32))
    target = rng.standard_normal((n, nx * ny, 2)).astype(np.float32)
    return feats, target


def test_forward_matches_oracle():
    net = learned.UNet(seed=3)
    theta = ur.bf16_round(net.get_params())
    net.set_params(theta)
    for nx, ny in [(17, 9), (33, 25)]:
        feats, _ = _case(nx, nx=nx, ny=ny)
        y = net.forward(feats, nx, ny)
        yr = ur.forward(theta.astype(np.float64), feats.astype(np.float64), nx, ny)
        err = np.linalg.norm(y - yr) / np.linalg.norm(yr)
        print(f"forward {nx}x{ny}: rel err {err:.2e}")
        assert err <= 3e-2


def test_gradient_matches_autograd():
    pytest.importorskip("torch")
    net = learned.UNet(seed=5)
    theta = ur.bf16_round(net.get_params())
    net.set_params(theta)
    feats, target = _case(9, n=2, nx=17, ny=17)
    loss, g = net.loss_grad(feats, target, 17, 17)
    lref, gref = torch_loss_grad(theta.astype(np.float64), feats.astype(np.float64), target.astype(np.float64), 17, 17)
    print(f"loss gpu {loss:.6f} ref {lref:.6f}")
    assert loss == pytest.approx(lref, rel=2e-2)
    err = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f"gradient rel err {err:.2e}")
    assert err <= 6e-2
    off = 0
    for l, (ci, co) in enumerate(ur.SPEC):
        n = 9 * co * ci + co
        a, b = g[off:off + n], gref[off:off + n]
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
        assert cos >= 0.99, (l, cos)
        off += n


def test_gradient_across_pixel_chunks_matches_autograd():
    """Three 65 x 33 images (2145 pixels each): the weight-gradient GEMM's
    64-pixel chunks and its pixel blocks straddle image boundaries, and every
    layer's gradient (all channel counts, every tap) must still match float64
    autograd of the independent torch statement."""
    pytest.importorskip("torch")
    net = learned.UNet(seed=6)
    theta = ur.bf16_round(net.get_params())
    net.set_params(theta)
    feats, target = _case(10, n=3, nx=65, ny=33)
    loss, g = net.loss_grad(feats, target, 65, 33)
    lref, gref = torch_loss_grad(theta.astype(np.float64), feats.astype(np.float64), target.astype(np.float64), 65, 33)
    assert loss == pytest.approx(lref, rel=2e-2)
    err = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f"gradient (3 x 65 x 33) rel err {err:.2e}")
    assert err <= 6e-2
    off = 0
    for l, (ci, co) in enumerate(ur.SPEC):
        nw = 9 * co * ci
        for part, (a, b) in (("weights", (g[off:off + nw], gref[off:off + nw])),
                             ("bias", (g[off + nw:off + nw + co], gref[off + nw:off + nw + co]))):
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
            assert cos >= 0.99, (l, part, cos)
        off += nw + co


def test_adam_step_closed_form():
    net = learned.UNet(seed=1)
    theta = net.get_params().astype(np.float64)
    g = np.random.default_rng(2).standard_normal(theta.size) * 1e-2
    net.adam_step(g.astype(np.float32), lr=1e-3)
    t1, m, v = ur.adam_step(theta, g.astype(np.float32).astype(np.float64), 0 * theta, 0 * theta, 1, lr=1e-3)
    np.testing.assert_allclose(net.get_params(), t1, rtol=1e-5, atol=1e-7)
    net.adam_step(g.astype(np.float32), lr=1e-3)
    t2, _, _ = ur.adam_step(t1, g.astype(np.float32).astype(np.float64), m, v, 2, lr=1e-3)
    np.testing.assert_allclose(net.get_params(), t2, rtol=1e-5, atol=1e-7)


def _problem(n, seed, ppw=8.0, h=10.0):
    v = media.layered_body_velocity_2d(n, seed=seed)
    m = (1.0 / v**2).ravel(order="F")
    M = ha.getRegularMesh([0.0, h * (n[0] - 1), 0.0, h * (n[1] - 1)], [n[0] - 1, n[1] - 1])
    w = media.omega_for_ppw(1500.0, h, ppw)
    _, gamma = ha.GetHelmholtzOperator(M, m, w, 0.01 * w * np.ones(m.size), True, [16, 16], w, True)
    return M, ha.HelmholtzParam(M, gamma.ravel(order="F"), m, w, True, True)


def _solver(hp, levels=5, basis="c128"):
    mg = ha.getMGparam(complex, int, levels, 1, 100, 1e-6, "Jac", 0.8, 2, 2, "W", "NoMUMPS", 0.5, 0.0)
    mg.prec = "c64"
    mg.basis = basis
    return ha.getShiftedLaplacianMultigridSolver(hp, mg, 0.5, "GMRES", 10)


def _sources(M, count, seed):
    rng = np.random.default_rng(seed)
    n1 = int(M.nodes[0])
    subs = [(int(x), 2) for x in rng.integers(20, n1 - 20, size=count)]
    return ha.point_sources(M, subs)


def test_training_lowers_loss_and_learned_solve_converges():
    n = (257, 257)
    level = 2                      # 65 x 65 nodes
    nl = (n[0] - 1) // 4 + 1
    data = learned.Dataset(nl, nl, 2048)
    for k in range(4):
        M, hp = _problem(n, seed=100 + k, ppw=6.0 + 2.0 * k)
        s = _solver(hp)
        learned.capture_begin(s, level, 512, 0, 8)
        ha.solveLinearSystem(None, _sources(M, 32, k), s)
        data.add_from_solver(s, sweeps=12)
    assert data.size() >= 512
    net = learned.UNet(seed=0)
    e0 = net.eval(data)
    losses = net.train(data, epochs=6, batch=16, lr=2e-3, seed=1)
    e1 = net.eval(data)
    print(f"losses per epoch {np.round(losses, 4)}; rel err of the correction {e0:.3f} -> {e1:.3f}")
    assert losses[-1] < losses[0]
    assert e1 < e0
    # evaluation on an unseen medium: plain W-cycle vs learned coarse correction
    M, hp = _problem(n, seed=999, ppw=8.0)
    B = _sources(M, 16, 77)
    H = hr.get_helmholtz_operator7(hr.get_regular_mesh(M.domain, M.n), hp.m, hp.omega, hp.gamma, True, True)
    sp = _solver(hp)
    Xp, sp = ha.solveLinearSystem(None, B, sp)
    sl = _solver(hp)
    learned.attach(sl, net, level)
    Xl, sl = ha.solveLinearSystem(None, B, sl)
    for X in (Xp, Xl):
        for c in range(B.shape[1]):
            assert np.linalg.norm(H @ X[:, c] - B[:, c]) / np.linalg.norm(B[:, c]) <= 1e-6
    print(f"ablation 257x257 8 ppw: plain W iters {sp.last_stats['max_iters']} "
          f"({sp.last_stats['solve_time']:.3f} s), learned iters {sl.last_stats['max_iters']} "
          f"({sl.last_stats['solve_time']:.3f} s), fell back: {sl.last_stats['fell_back']}")
    # the restart-0/restart-1 trial bounds what a weak network can cost: two
    # restart cycles of 10 iterations at most
    assert sl.last_stats["max_iters"] <= sp.last_stats["max_iters"] + 20
    # the verdict is kept: a second solve on the same hierarchy skips the trial
    first_fell_back = sl.last_stats["fell_back"]
    Xl2, sl = ha.solveLinearSystem(None, B, sl)
    if first_fell_back:
        assert sl.last_stats["fell_back"] and not sl.last_stats["learned_used"]
        assert sl.last_stats["max_iters"] == sp.last_stats["max_iters"]
        assert np.array_equal(Xl2, Xp)
    else:
        assert sl.last_stats["learned_used"]


def test_learned_trial_with_c64_basis():
    """The learned correction's restart (normalised complex64 inputs scaled
    from the complex64 basis) and the trial against the plain cycle: every
    RHS reaches the 1e-6 true residual from the oracle's CSR operator with an
    untrained network attached, complex64 and complex128 basis alike."""
    n = (257, 257)
    M, hp = _problem(n, seed=321, ppw=8.0)
    B = _sources(M, 8, 5)
    H = hr.get_helmholtz_operator7(hr.get_regular_mesh(M.domain, M.n), hp.m, hp.omega, hp.gamma, True, True)
    net = learned.UNet(seed=0)
    for basis in ("c128", "c64"):
        s = _solver(hp, basis=basis)
        learned.attach(s, net, 2)
        X, s = ha.solveLinearSystem(None, B, s)
        res = np.linalg.norm(H @ X - B, axis=0) / np.linalg.norm(B, axis=0)
        print(f"{basis} basis, untrained network: iters {s.last_stats['max_iters']}, "
              f"fell back {s.last_stats['fell_back']}, max true residual {res.max():.2e}")
        assert res.max() <= 1e-6
