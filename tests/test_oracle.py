"""CPU pins of the oracle (no GPU).

Helmholtz.jl's tests hold no assertions or golden vectors
(test/testFictitiousSource2D.jl:57 is the only @test, commented out), so the
oracle is pinned by: explicit KATs typed from the reference's row formulas,
independent re-derivations (mirror-ghost loops, the reference's own
matrix-free ``multOpNeumann!``), closed forms, invariants, the reference's
manufactured-solution flow (second-order convergence) and committed fixtures.
Each check is chosen so that a plausible mistake (a dropped term, a wrong
sign, a transposed kron, an off-by-one pad) fails it.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import helmholtz_ref as hr, slmg_ref as sr
from tests.problems import attenuation_case, attenuation_decay_slope, seg_salt_case, small_case

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- Laplacian
def test_dxx_mat_kat_bc2():
    # src/PlainNodalLaplacian.jl:18-30 rows [BC,-BC], [-1,2,-1], [-BC,BC] over h^2, BC=2
    D = hr.dxx_mat(4, 0.5, 2).toarray()
    K = np.array([[2, -2, 0, 0], [-1, 2, -1, 0], [0, -1, 2, -1], [0, 0, -2, 2]]) / 0.25
    np.testing.assert_array_equal(D, K)


def test_dxx_mat_kat_bc1():
    D = hr.dxx_mat(4, 0.5, 1).toarray()
    K = np.array([[1, -1, 0, 0], [-1, 2, -1, 0], [0, -1, 2, -1], [0, 0, -1, 1]]) / 0.25
    np.testing.assert_array_equal(D, K)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("cells", [[6, 5], [4, 3, 5]])
def test_laplacian_kills_constants(order, cells):
    M = hr.get_regular_mesh(sum(([0, 1 + d] for d in range(len(cells))), []), cells)
    L = hr.get_nodal_laplacian_matrix(M, order)
    assert np.abs(L @ np.ones(L.shape[0])).max() <= 1e-14 * abs(L).max()


def _mirror_lap(X, h, mode):
    """Independent statement of the nodal Laplacian: pad every axis by one
    ghost (mirror x~_{-1} = x_1 for BC=2, copy x~_{-1} = x_0 for BC=1) and
    apply the (2x - left - right)/h^2 formula per axis."""
    pad_mode = "reflect" if mode == 2 else "edge"
    G = np.pad(X, 1, mode=pad_mode)
    out = np.zeros_like(X)
    core = tuple(slice(1, -1) for _ in X.shape)
    for d in range(X.ndim):
        lo = list(core)
        hi = list(core)
        lo[d] = slice(0, -2)
        hi[d] = slice(2, None)
        out += (2 * G[core] - G[tuple(lo)] - G[tuple(hi)]) / h[d] ** 2
    return out


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("cells", [[7, 5], [5, 4, 6]])
def test_kron_laplacian_equals_ghost_stencil(order, cells):
    M = hr.get_regular_mesh(sum(([0, 1 + 0.3 * d] for d in range(len(cells))), []), cells)
    L = hr.get_nodal_laplacian_matrix(M, order)
    nodes = tuple(int(v) for v in M.nodes)
    x = np.random.default_rng(0).standard_normal(int(np.prod(nodes)))
    ref = _mirror_lap(x.reshape(nodes, order="F"), M.h, order).ravel(order="F")
    np.testing.assert_allclose(L @ x, ref, rtol=1e-13, atol=1e-10)


def test_bc1_equals_reference_multOpNeumann():
    # src/PlainNodalLaplacian.jl:155-188 (Lap2DStencil with ghost = self)
    M = hr.get_regular_mesh([0, 2, 0, 1.5], [9, 6])
    x = np.random.default_rng(1).standard_normal(int(np.prod(M.nodes)))
    L = hr.get_nodal_laplacian_matrix(M, 1)
    np.testing.assert_allclose(L @ x, hr.mult_op_neumann(M, x), rtol=1e-13, atol=1e-10)


def test_bc2_is_not_symmetric_but_bc1_is():
    # comment at src/PlainNodalLaplacian.jl:7: with 2 the matrix is not symmetric
    M = hr.get_regular_mesh([0, 1, 0, 1], [4, 4])
    L2 = hr.get_nodal_laplacian_matrix(M, 2)
    L1 = hr.get_nodal_laplacian_matrix(M, 1)
    assert abs(L2 - L2.T).max() > 0
    assert abs(L1 - L1.T).max() == 0


# ---------------------------------------------------------------- ABL
def test_abl_2d_closed_form():
    n, pad, A = [12, 10], [3, 4], 2.0
    g = hr.get_abl(n, False, pad, A)
    # corner: both ramps at full strength -> A (1 + 1 - 1)
    assert g[0, 0] == pytest.approx(A)
    # a point only in the left x ramp at k = 1 (0-based): ((pad - 1)/pad)^2
    assert g[1, 5] == pytest.approx(A * (2 / 3) ** 2)
    # right x ramp at the last node: (pad/pad)^2 = 1 ; right ramp first node (1/pad)^2
    assert g[11, 5] == pytest.approx(A)
    assert g[9, 5] == pytest.approx(A * (1 / 3) ** 2)
    # interior untouched
    assert g[5, 5] == 0.0
    # mixed: p1 = (1/3)^2 at i=9, p2 = (2/4)^2 at j=8 (high z ramp k=2 -> (3/4)^2?)
    p1 = (1 / 3) ** 2
    p2 = (3 / 4) ** 2  # j = 8 is the 3rd node of the last-4 ramp -> (3/4)^2
    assert g[9, 8] == pytest.approx(A * (p1 + p2 - p1 * p2))


def test_abl_2d_neumann_top_has_no_top_ramp():
    g = hr.get_abl([12, 10], True, [3, 4], 1.0)
    assert g[5, 0] == 0.0
    assert g[0, 0] == pytest.approx(1.0)  # left ramp only
    g2 = hr.get_abl([12, 10], False, [3, 4], 1.0)
    assert g2[5, 0] == pytest.approx(1.0)


def test_abl_3d_closed_form():
    n, pad, A = [9, 8, 7], [2, 3, 2], 3.0
    g = hr.get_abl(n, False, pad, A)
    x = np.linspace(-1, 1, 9)
    r0 = (x[0] - x[1]) ** 2  # left ramp at i = 0 (pad 2: reference node x[pad-1])
    gx0 = r0 / (r0 + 1e-5)  # both x ramps peak at the same value
    # a node in the x ramp only (y, z interior)
    assert g[0, 4, 3] == pytest.approx(A * gx0)
    assert g[4, 4, 3] == 0.0
    assert g.max() <= A
    assert g[0, 0, 0] == A  # clipped


# ---------------------------------------------------------------- Sommerfeld and operator
def test_sommerfeld_closed_form():
    M = hr.get_regular_mesh([0, 2, 0, 1], [8, 4])
    m = np.linspace(0.2, 0.5, 45)
    w = 3.0
    S = hr.get_sommerfeld_bc(M, m, w, True)
    mm = m.reshape((9, 5), order="F")
    h1, h2 = M.h
    assert S[4, 2] == 0
    assert S[8, 2] == pytest.approx(-1j * w * 2 / h1 * np.sqrt(mm[8, 2]))
    assert S[8, 4] == pytest.approx(-1j * w * (2 / h1 + 2 / h2) * np.sqrt(mm[8, 4]))
    assert S[4, 0] == 0  # Neumann top: no top term
    S2 = hr.get_sommerfeld_bc(M, m, w, False)
    assert S2[4, 0] == pytest.approx(-1j * w * 2 / h2 * np.sqrt(mm[4, 0]))


def test_operator_diagonal_closed_form():
    M = hr.get_regular_mesh([0, 1, 0, 1], [6, 6])
    N = 49
    m = np.full(N, 0.25)
    g = np.full(N, 0.7)
    w = 5.0
    H = hr.get_helmholtz_operator7(M, m, w, g, False, False)
    h = M.h[0]
    # interior node: 4/h^2 - w^2 m (1 - i g / w)
    assert H[24, 24] == pytest.approx(4 / h**2 - w**2 * 0.25 * (1 - 1j * 0.7 / w))
    assert H[24, 23] == pytest.approx(-1 / h**2)


def test_operator_real_without_attenuation():
    M = hr.get_regular_mesh([0, 1, 0, 1], [8, 6])
    m = np.random.default_rng(0).uniform(0.1, 0.4, 63)
    H = hr.get_helmholtz_operator7(M, m, 4.0, np.zeros(63), False, False)
    assert abs(H.imag).max() == 0.0


def test_shifted_param_equals_shift_op():
    M = hr.get_regular_mesh([0, 1, 0, 2], [10, 8])
    m = np.random.default_rng(2).uniform(0.1, 0.4, 99)
    g = np.random.default_rng(3).uniform(0.0, 2.0, 99)
    p = hr.HelmholtzParam(M, g, m, 6.0, True, True)
    H = hr.operator_from_param(p)
    SH = hr.operator_from_param(hr.get_shifted_helmholtz_param(p, 0.3))
    D = SH - (H + hr.get_helmholtz_shift_op(m, 6.0, 0.3))
    assert abs(D).max() <= 1e-15 * abs(SH).max()


def test_helmholtz_fun_recovers_h():
    # getHelmholtzFun(SH^H, ShiftOP) applies SH x - i b w^2 m x = H x
    M = hr.get_regular_mesh([0, 1, 0, 1], [8, 8])
    m = np.random.default_rng(4).uniform(0.1, 0.4, 81)
    H = hr.get_helmholtz_operator7(M, m, 5.0, np.full(81, 0.3), False, True)
    S = hr.get_helmholtz_shift_op(m, 5.0, 0.5)
    SHt = (H + S).conj().T  # the solver receives sparse(SH') (test/ShiftedLaplacianTest.jl:83)
    x = np.random.default_rng(5).standard_normal(81) + 1j
    # with SpMatMul's adjoint semantics the second product is (ShiftOP)^H x = -i b w^2 m x
    f2 = hr.helmholtz_fun(SHt, S)
    np.testing.assert_allclose(f2(x), H @ x, rtol=1e-13, atol=1e-12)


def test_max_frequency_is_ten_ppw():
    M = hr.get_regular_mesh([0, 1, 0, 1], [100, 100])
    m = np.array([1 / 2.0**2, 1 / 3.0**2])
    w = hr.get_maximal_frequency(m, M)
    lam = 2 * np.pi * 2.0 / w  # wavelength at v_min = 2
    assert lam / M.h[0] == pytest.approx(10.0)


def test_point_source_and_index():
    M = hr.get_regular_mesh([0, 1, 0, 1], [128, 128])
    q, src = hr.get_acoustic_point_source(M)
    assert src == [64, 1]  # div(n+1, 2), top row
    assert hr.loc2cs([129, 129], [64, 1]) == 64
    assert q[63] == pytest.approx(1 / np.linalg.norm(M.h) ** 2)
    assert hr.get_mid_point_src(M) == [64, 64]
    assert hr.loc2cs([5, 4, 3], [2, 3, 2]) == 2 + 2 * 5 + 1 * 20


# ---------------------------------------------------------------- manufactured solution
def _fictitious_source_errors(slowness, refinements=3):
    """test/testFictitiousSource2D.jl:10-55 flow on [-1,1]^2 from 32x48 cells,
    u = cos(pi x) cos(pi y) (zero normal derivative), Neumann rows, ABL gamma
    (pad 5, amplitude 0.1) with a right-hand side consistent with the
    operator's -w^2 m (1 - i gamma / w) mass term."""
    w = 2 * np.pi
    errs = []
    n = np.array([32, 48])
    for _ in range(refinements):
        n = n * 2
        M = hr.get_regular_mesh([-1, 1, -1, 1], n)
        x = np.linspace(-1, 1, n[0] + 1)
        y = np.linspace(-1, 1, n[1] + 1)
        X, Y = np.meshgrid(x, y, indexing="ij")
        u = (np.cos(np.pi * X) * np.cos(np.pi * Y)).ravel(order="F")
        m = slowness(X, Y).ravel(order="F")
        g = hr.get_abl(M.nodes, True, [5, 5], 0.1).ravel(order="F")
        f = 2 * np.pi**2 * u - w**2 * m * (1 - 1j * g / w) * u
        A = hr.get_helmholtz_operator7(M, m, w, g, True, False)
        ut = spla.spsolve(A.tocsc(), f)
        errs.append(np.abs(ut - u).max())
    return np.array(errs)


@pytest.mark.parametrize("slowness", [lambda X, Y: np.ones_like(X),
                                      lambda X, Y: np.exp(-2.0 * (X**2 + Y**2))])
def test_second_order_convergence(slowness):
    e = _fictitious_source_errors(slowness)
    ratios = e[:-1] / e[1:]
    assert np.all(ratios > 3.6) and np.all(ratios < 4.4), (e, ratios)


# ---------------------------------------------------------------- multigrid / FGMRES invariants
def test_prolongation_preserves_linears():
    P = sr.prolongation([5, 4])
    xc = np.linspace(0, 1, 5)
    yc = np.linspace(0, 2, 4)
    Xc, Yc = np.meshgrid(xc, yc, indexing="ij")
    xf = np.linspace(0, 1, 9)
    yf = np.linspace(0, 2, 7)
    Xf, Yf = np.meshgrid(xf, yf, indexing="ij")
    lin_c = (3 * Xc - 2 * Yc + 1).ravel(order="F")
    lin_f = (3 * Xf - 2 * Yf + 1).ravel(order="F")
    np.testing.assert_allclose(P @ lin_c, lin_f, atol=1e-14)


@pytest.mark.parametrize("nodes", [[5, 4], [3, 4, 5]])
def test_restriction_is_weighted_transpose(nodes):
    R = sr.restriction(nodes).toarray()
    P = sr.prolongation(nodes).toarray()
    np.testing.assert_allclose(R @ np.ones(R.shape[1]), 1.0, atol=1e-15)

    def trap(n):
        w = np.ones(n)
        w[0] = w[-1] = 0.5
        return w

    Wc = np.ones(1)
    Wf = np.ones(1)
    for n in nodes:
        Wc = np.kron(trap(n), Wc)
        Wf = np.kron(trap(2 * n - 1), Wf)
    np.testing.assert_allclose(R, (P.T * Wf[None, :]) / Wc[:, None] / 2 ** len(nodes), atol=1e-15)


def _small_problem(cells=(64, 64), w_scale=1.0):
    M = hr.get_regular_mesh([0, 1, 0, 1], list(cells))
    nodes = M.nodes
    m = np.random.default_rng(0).uniform(0.2, 0.3, int(np.prod(nodes)))
    w = hr.get_maximal_frequency(m, M) * w_scale
    g = hr.get_abl(nodes, False, [6, 6], w).ravel(order="F") + 0.01 * w
    p = hr.HelmholtzParam(M, g, m, w, False, True)
    return p, hr.operator_from_param(p)


def test_vcycle_contracts_on_shifted_operator():
    p, H = _small_problem()
    hier = sr.build_hierarchy(p, 3, 0.5)
    SH = hier[0].A
    b = np.random.default_rng(1).standard_normal(SH.shape[0]) + 0j
    x = np.zeros_like(b)
    r0 = np.linalg.norm(b)
    for _ in range(6):
        x = x + sr.vcycle(hier, 0, b - SH @ x)
    assert np.linalg.norm(b - SH @ x) / r0 < 0.5**6


def test_fgmres_exact_preconditioner_one_iteration():
    p, H = _small_problem((32, 32))
    lu = spla.splu(H.tocsc())
    b = np.random.default_rng(2).standard_normal(H.shape[0]) + 0j
    x, it, hist = sr.fgmres(lambda v: H @ v, b, lu.solve, 10, 5, 1e-10)
    assert it == 1
    assert np.linalg.norm(H @ x - b) / np.linalg.norm(b) < 1e-10


def test_fgmres_estimate_equals_true_residual_and_is_monotone():
    p, H = _small_problem((32, 32))
    hier = sr.build_hierarchy(p, 2, 0.5)
    b = np.random.default_rng(3).standard_normal(H.shape[0]) + 1j
    x, it, hist = sr.fgmres(lambda v: H @ v, b, lambda v: sr.vcycle(hier, 0, v), 40, 3, 1e-9)
    assert np.all(np.diff(hist[: it + 1]) <= 1e-15)
    true = np.linalg.norm(H @ x - b) / np.linalg.norm(b)
    assert true <= 1e-9 and abs(true - hist[-1]) < 1e-11


def test_fgmres_identity_preconditioner_matches_direct():
    M = hr.get_regular_mesh([0, 1, 0, 1], [12, 12])
    m = np.full(169, 0.1)
    H = hr.get_helmholtz_operator7(M, m, 1.0, np.full(169, 0.5), False, True)
    b = np.random.default_rng(4).standard_normal(169) + 0j
    x, it, _ = sr.fgmres(lambda v: H @ v, b, lambda v: v, 30, 50, 1e-12)
    np.testing.assert_allclose(x, spla.spsolve(H.tocsc(), b), rtol=1e-9, atol=1e-12)


def test_golden_fixtures_unchanged():
    """Committed operator/ABL fixtures (tests/golden/make_golden.py) still
    match the oracle bit for bit (guards against silent oracle drift)."""
    f = np.load(os.path.join(GOLDEN, "operator_fixtures.npz"))
    for key in sorted({k.split("__")[0] for k in f.files}):
        spec = f[key + "__spec"]
        dim = int(spec[0])
        nodes = [int(v) for v in spec[1:1 + dim]]
        neumann, somm, order = bool(spec[4]), bool(spec[5]), int(spec[6])
        M = hr.get_regular_mesh(f[key + "__domain"], [n - 1 for n in nodes])
        m, g, w = f[key + "__m"], f[key + "__gamma"], complex(f[key + "__omega"])
        H = hr.get_helmholtz_operator7(M, m, w, g, neumann, somm, order)
        np.testing.assert_array_equal(H @ f[key + "__x"], f[key + "__y"])


def test_bicgstab_exact_preconditioner_one_iteration():
    M, m, g, w = small_case((17, 9), seed=5)
    H = hr.get_helmholtz_operator7(M, m, w, g, True, True).tocsc()
    lu = spla.splu(H)
    b = np.zeros(m.size, dtype=complex)
    b[20] = 1.0
    x, it, hist = sr.bicgstab(lambda v: H @ v, b, lu.solve, 10, 1e-10)
    assert it == 1
    assert np.linalg.norm(H @ x - b) <= 1e-12 * np.linalg.norm(b)


def test_bicgstab_identity_preconditioner_matches_direct():
    rng = np.random.default_rng(3)
    n = 40
    A = np.eye(n) * 4 + rng.standard_normal((n, n)) * 0.3 + 1j * rng.standard_normal((n, n)) * 0.3
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x, it, hist = sr.bicgstab(lambda v: A @ v, b, lambda v: v, 200, 1e-12)
    assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-10 * np.linalg.norm(x)
    assert hist[-1] <= 1e-12


@pytest.mark.parametrize("nodes,order", [((9, 7), 2), ((9, 7), 1), ((5, 6, 4), 2), ((5, 6, 4), 1)])
def test_adjoint_similarity_with_trapezoid_weights(nodes, order):
    """H^H = W conj(H) W^-1 on the assembled operator: the identity the
    adjoint (doTranspose) solve is built on."""
    M, m, g, w = small_case(nodes, seed=8)
    H = hr.get_helmholtz_operator7(M, m, w, g, True, True, order).toarray()
    W = np.diag(sr.trapezoid_weights(M, order))
    assert np.allclose(H.conj().T, W @ H.conj() @ np.linalg.inv(W), rtol=0, atol=1e-12 * np.abs(H).max())
    if order == 2:
        assert not np.allclose(H.T, H)  # the identity is not a triviality


@pytest.mark.parametrize("nodes", [(9, 7), (6, 5, 4)])
def test_ho_beta_one_is_the_bc1_operator(nodes):
    """GetHelmholtzOperatorHO with beta = 1: G'G is the natural-Neumann
    Laplacian and av3term(., 1) = I, so H_HO equals the 7-argument operator
    with orderNeumannBC = 1 (src/GetHelmholtz.jl:33-50 vs 54-72)."""
    M, m, g, w = small_case(nodes, seed=9)
    H1 = hr.get_helmholtz_operator_ho(M, m, w, g, True, True, 1.0)
    H7 = hr.get_helmholtz_operator7(M, m, w, g, True, True, 1)
    assert abs(H1 - H7).max() <= 1e-13 * abs(H7).max()


@pytest.mark.parametrize("nodes,beta", [((9, 7), 0.3), ((9, 7), 2.0 / 3.0), ((6, 5, 4), (0.5, 0.25))])
def test_ho_laplacian_and_mass_invariants(nodes, beta):
    """Spread Laplacian: symmetric, kills constants, 9-point (2-D) / 19-point
    (3-D) interior rows; averaging mass rows sum to 1."""
    M, m, g, w = small_case(nodes, seed=10)
    Lap, Mm = hr.get_spread_nodal_laplacian_and_mass(M, beta)
    N = Lap.shape[0]
    assert abs(Lap - Lap.T).max() <= 1e-12 * abs(Lap).max()
    assert np.abs(Lap @ np.ones(N)).max() <= 1e-10 * abs(Lap).max()
    assert np.allclose(Mm @ np.ones(N), 1.0, atol=1e-14)
    counts = np.diff(Lap.tocsr().indptr)
    assert counts.max() == (9 if M.dim == 2 else 19)
    # the mass term multiplies before it averages: H = Lap + M diag(mass)
    H = hr.get_helmholtz_operator_ho(M, m, w, g, True, True, beta)
    mass = hr.helmholtz_mass(M, m, w, g, True, True)
    x = np.random.default_rng(1).standard_normal(N) + 0j
    assert np.allclose(H @ x, Lap @ x + Mm @ (mass * x), rtol=1e-13, atol=0)
    # the averaging mass matrix is symmetric, so the transpose the adjoint solve
    # applies is H^T = Lap + diag(mass) M; H itself is not symmetric (the
    # diagonal varies with the medium), so the two orders differ
    assert abs(Mm - Mm.T).max() <= 1e-15
    assert np.allclose(H.T @ x, Lap @ x + mass * (Mm @ x), rtol=1e-13, atol=0)
    assert abs(H - H.T).max() > 1e-6 * abs(H).max()


def test_kcycle_correction_is_minimal_residual():
    """The K-cycle's coarse correction e minimises ||rc - A e|| over span{c1, c2}
    (the two coarse cycles it applies): the residual is orthogonal to A c1 and
    A c2, and random combinations do no better."""
    M, m, g, w = small_case((65, 33), seed=11)
    p = hr.HelmholtzParam(M, g, m, w, True, True)
    hier = sr.build_hierarchy(p, 4, 0.5)
    rng = np.random.default_rng(5)
    rc = rng.standard_normal(hier[1].A.shape[0]) + 1j * rng.standard_normal(hier[1].A.shape[0])
    e = sr.kcycle_correction(hier, 0, rc, 0.8, 2, 2, 0)
    A = hier[1].A
    c1 = sr.vcycle(hier, 1, rc, 0.8, 2, 2, "K")
    v1 = A @ c1
    a1 = np.vdot(v1, rc) / np.vdot(v1, v1).real
    c2 = sr.vcycle(hier, 1, rc - a1 * v1, 0.8, 2, 2, "K")
    res = rc - A @ e
    for c in (c1, c2):
        Ac = A @ c
        assert abs(np.vdot(Ac, res)) <= 1e-10 * np.linalg.norm(Ac) * np.linalg.norm(res)
    for _ in range(20):
        al, be = rng.standard_normal(2) + 1j * rng.standard_normal(2)
        assert np.linalg.norm(res) <= np.linalg.norm(rc - A @ (al * c1 + be * c2)) + 1e-12


def test_kcycle_converges_like_w_and_beats_v():
    M, m, g, w = small_case((129, 65), seed=12)
    p = hr.HelmholtzParam(M, g, m, w, True, True)
    H = hr.operator_from_param(p)
    hier = sr.build_hierarchy(p, 4, 0.5)
    b = np.zeros(m.size, complex)
    b[64 + 129 * 2] = 1.0
    its = {}
    for cyc in ("V", "W", "K"):
        x, its[cyc], _ = sr.fgmres(lambda v: H @ v, b, lambda v: sr.vcycle(hier, 0, v, 0.8, 2, 2, cyc), 10, 80, 1e-6)
        assert np.linalg.norm(H @ x - b) <= 1e-6 * np.linalg.norm(b)
    assert its["K"] < its["V"]


@pytest.mark.parametrize("cycle", ["V", "W"])
def test_per_level_sweeps_equal_error_propagation(cycle):
    """Per-level sweep counts (relaxPre/relaxPost as functions of the level):
    the cycle's action equals (I - E_0) A_0^-1 built from the error-propagation
    matrices E_l = S_l^post_l (I - P_l B_{l+1} R_l A_l)^k S_l^pre_l, with
    S_l = I - w D_l^-1 A_l, B_{l+1} = (I - E_{l+1}) A_{l+1}^-1, B_last = A_last^-1
    and k = 1 (V) or 2 (W).  Unequal pre/post counts make a swapped pair fail."""
    M, m, g, w = small_case((17, 13), seed=31)
    p = hr.HelmholtzParam(M, g, m, w, True, True)
    hier = sr.build_hierarchy(p, 3, 0.5)
    pre, post = (lambda l: 2 * l + 1), (lambda l: 4 - l)
    k = 2 if cycle == "W" else 1
    B = np.linalg.inv(hier[-1].A.toarray())
    for l in range(len(hier) - 2, -1, -1):
        lev = hier[l]
        A = lev.A.toarray()
        I = np.eye(A.shape[0])
        S = I - 0.8 * lev.Dinv[:, None] * A
        T = I - lev.P.toarray() @ B @ lev.R.toarray() @ A
        E = (np.linalg.matrix_power(S, post(l + 1)) @ np.linalg.matrix_power(T, k)
             @ np.linalg.matrix_power(S, pre(l + 1)))
        B = (I - E) @ np.linalg.inv(A)
    rng = np.random.default_rng(32)
    b = rng.standard_normal(m.size) + 1j * rng.standard_normal(m.size)
    z = sr.vcycle(hier, 0, b, 0.8, pre, post, cycle)
    assert np.linalg.norm(z - B @ b) <= 1e-10 * np.linalg.norm(z)
    # the same counts as sequences indexed by level
    z2 = sr.vcycle(hier, 0, b, 0.8, [3, 5, 7], [3, 2, 1], cycle)
    assert np.array_equal(z, z2)
    # the swapped pair is a different cycle
    zs = sr.vcycle(hier, 0, b, 0.8, post, pre, cycle)
    assert np.linalg.norm(zs - z) > 1e-6 * np.linalg.norm(z)


@pytest.mark.parametrize("steps", [1, 2, 3])
def test_jac_gmres_is_minimal_residual_over_preconditioned_krylov_space(steps):
    """Jac-GMRES smoothing: the new iterate minimises ||b - A x|| over
    x0 + M span{r0, (A M) r0, ..., (A M)^(s-1) r0} with M = w D^-1.  Checked
    against a brute-force least-squares solve on the explicit Krylov basis;
    one step has the closed form alpha = <A M r0, r0> / ||A M r0||^2."""
    M, m, g, w = small_case((17, 13), seed=41)
    p = hr.HelmholtzParam(M, g, m, w, True, True)
    lev = sr.build_hierarchy(p, 2, 0.5)[0]
    rng = np.random.default_rng(42)
    n = lev.A.shape[0]
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Mi = 0.8 * lev.Dinv
    r0 = b - lev.A @ x0
    K = [Mi * r0]
    for _ in range(steps - 1):
        K.append(Mi * (lev.A @ K[-1]))
    K = np.stack(K, axis=1)
    c = np.linalg.lstsq(lev.A @ K, r0, rcond=None)[0]
    xb = x0 + K @ c
    x = sr.jac_gmres(lev, x0, b, 0.8, steps)
    rb, rx = np.linalg.norm(b - lev.A @ xb), np.linalg.norm(b - lev.A @ x)
    assert rx <= rb * (1 + 1e-9)
    assert np.linalg.norm(x - xb) <= 1e-8 * np.linalg.norm(xb)
    if steps == 1:
        q = lev.A @ (Mi * r0)
        alpha = np.vdot(q, r0) / np.vdot(q, q)
        assert np.linalg.norm(x - (x0 + alpha * Mi * r0)) <= 1e-12 * np.linalg.norm(x)
    # the smoother never increases the residual and zero steps change nothing
    assert rx <= np.linalg.norm(r0)
    assert np.array_equal(sr.jac_gmres(lev, x0, b, 0.8, 0), x0)


def test_kcycle_with_jac_gmres_converges():
    """The reference test's last configuration (test/ShiftedLaplacianTest.jl:137-142):
    K-cycle with Jac-GMRES relaxation inside FGMRES(5) reaches 1e-6."""
    M, m, g, w = small_case((65, 65), seed=43)
    p = hr.HelmholtzParam(M, g, m, w, True, True)
    H = hr.operator_from_param(p)
    hier = sr.build_hierarchy(p, 3, 0.5)
    b = np.zeros(m.size, complex)
    b[32 + 65 * 32] = 1.0
    x, its, _ = sr.fgmres(lambda v: H @ v, b, lambda v: sr.vcycle(hier, 0, v, 0.8, 2, 2, "K", 0, "Jac-GMRES"),
                          5, 60, 1e-6)
    assert np.linalg.norm(H @ x - b) <= 1e-6 * np.linalg.norm(b)
    xj, itsj, _ = sr.fgmres(lambda v: H @ v, b, lambda v: sr.vcycle(hier, 0, v, 0.8, 2, 2, "K", 0, "Jac"), 5, 60, 1e-6)
    print(f"K-cycle FGMRES(5): Jac-GMRES {its} iterations, Jac {itsj}")


def test_attenuation_real_and_complex_frequency_agree():
    """test/AttenuationTest.jl: attenuation alpha added to gamma at real w and
    the complex frequency w - i alpha/2 give nearly the same field (they differ
    by the alpha^2/4 and alpha*gamma/2 terms of the mass, O((alpha/2w)^2)), and
    both decay at Im(kappa) = Im sqrt(w^2 + i w alpha) ~ alpha/2 (unit
    slowness).  A dropped attenuation term, a wrong sign of Im(w^2) or Re(w)
    in place of w in the mass each fail one of these."""
    M, m, w, alpha, gam, pad, src, q = attenuation_case()
    H1 = hr.get_helmholtz_operator7(M, m, w, gam + alpha, False, False)
    H2 = hr.get_helmholtz_operator7(M, m, w - 1j * alpha / 2, gam, False, False)
    s1 = spla.splu(H1.tocsc()).solve(q)
    s2 = spla.splu(H2.tocsc()).solve(q)
    rel = np.linalg.norm(s1 - s2) / np.linalg.norm(s1)
    kappa_im = np.sqrt(w**2 + 1j * w * alpha).imag
    nn = int(M.n[0]) + 1
    d1 = attenuation_decay_slope(s1, nn, src, pad, M.h[0])
    d2 = attenuation_decay_slope(s2, nn, src, pad, M.h[0])
    print(f"attenuation: rel diff {rel:.3e}; decay {d1:.4f} / {d2:.4f}, Im kappa {kappa_im:.4f}")
    assert rel <= 0.05
    assert abs(-d1 - kappa_im) <= 0.05 * kappa_im
    assert abs(-d2 - kappa_im) <= 0.05 * kappa_im
    # without the attenuation the field does not decay at that rate
    H0 = hr.get_helmholtz_operator7(M, m, w, gam, False, False)
    d0 = attenuation_decay_slope(spla.splu(H0.tocsc()).solve(q), nn, src, pad, M.h[0])
    assert -d0 < 0.25 * kappa_im


def test_seg_salt_fixture_and_reference_setup_converges():
    """The SEG salt slice (examples/SEGmodel2Dsalt.dat) loads with its
    checksum, water on top (1.5 km/s along the first depth row) and the salt
    body's 4.48 km/s inside; with the reference test's own solver settings
    (test/ShiftedLaplacianTest.jl:22-66: 2 levels, shift 0.02, W(2,2) Jacobi
    0.75, FGMRES(20)) the oracle's SL-MG reaches 1e-6 and the direct solve."""
    M, m, g, w, H, q = seg_salt_case()
    v = 1.0 / np.sqrt(m.reshape((257, 129), order="F"))
    assert np.allclose(v[:, 0], 1.5) and abs(v.max() - 4.48) < 1e-12 and abs(v.min() - 1.5) < 1e-12
    p = hr.HelmholtzParam(M, g, m, w, True, True)
    hier = sr.build_hierarchy(p, 2, 0.02)
    x, it, _ = sr.fgmres(lambda u: H @ u, q, lambda u: sr.vcycle(hier, 0, u, 0.75, 2, 2, "W"), 20, 30, 1e-6)
    res = np.linalg.norm(H @ x - q) / np.linalg.norm(q)
    xd = spla.splu(H.tocsc()).solve(q)
    err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
    print(f"SEG salt, reference settings: {it} iterations, res {res:.2e}, err vs direct {err:.2e}")
    assert res <= 1e-6 and it <= 30 and err <= 1e-4


def test_seg_fixture_matches_the_reference_file():
    """tests/golden/seg_salt_2d.npz holds exactly the values of the
    reference's examples/SEGmodel2Dsalt.dat (checked where the reference
    checkout is present; the GPU tests read only the fixture)."""
    src = "/root/reference/examples/SEGmodel2Dsalt.dat"
    if not os.path.exists(src):
        pytest.skip("reference checkout not present")
    f = np.load(os.path.join(GOLDEN, "seg_salt_2d.npz"))
    np.testing.assert_array_equal(f["v"].astype(np.float64), np.loadtxt(src))


def test_spread_operator_interior_weights_known_answer():
    """Interior 9-point row of the spread Laplacian G'((1-b) Gs + b G) and
    5-point row of the mass average M at b = 2/3, typed from the algebra of
    src/PlainNodalLaplacian.jl:59-66,106-120: Gs averages transversally with
    av3term(., 0.5) = [1/4, 1/2, 1/4], so the transverse weights are
    w(+-1) = (1 - b)/4 = 1/12 and w(0) = (1 - b)/2 + b = 5/6.  With
    A = 1/h1^2, B = 1/h2^2 the row is: centre (5/3)(A + B); x-neighbours
    -(5/6)A + (1/6)B; y-neighbours (1/6)A - (5/6)B; corners -(1/12)(A + B).
    M = (av3term_y(b) (x) I + I (x) av3term_x(b))/2: centre 2/3, the four
    neighbours 1/12.  (av3term's default weight 5/6 in Gs would make the
    corners -(1/36)(A + B).)"""
    mesh = hr.get_regular_mesh([0.0, 1.0, 0.0, 1.3], [8, 6])
    h1, h2 = mesh.h
    A, B = 1.0 / h1**2, 1.0 / h2**2
    Lap, M = hr.get_spread_nodal_laplacian_and_mass(mesh, 2.0 / 3.0)
    n1 = int(mesh.nodes[0])
    i, j = 4, 3  # 0-based interior node
    row = Lap.getrow(i + n1 * j).toarray().ravel()
    mrow = M.getrow(i + n1 * j).toarray().ravel()
    want = {(0, 0): 5.0 / 3.0 * (A + B), (1, 0): -5.0 / 6.0 * A + B / 6.0, (-1, 0): -5.0 / 6.0 * A + B / 6.0,
            (0, 1): A / 6.0 - 5.0 / 6.0 * B, (0, -1): A / 6.0 - 5.0 / 6.0 * B}
    for di in (-1, 1):
        for dj in (-1, 1):
            want[(di, dj)] = -(A + B) / 12.0
    got = {}
    for k in np.flatnonzero(row):
        got[(int(k % n1) - i, int(k // n1) - j)] = row[k]
    assert set(got) == set(want)
    for key, v in want.items():
        assert got[key] == pytest.approx(v, rel=1e-13), key
    mwant = {(0, 0): 2.0 / 3.0, (1, 0): 1.0 / 12.0, (-1, 0): 1.0 / 12.0, (0, 1): 1.0 / 12.0, (0, -1): 1.0 / 12.0}
    mgot = {(int(k % n1) - i, int(k // n1) - j): mrow[k] for k in np.flatnonzero(mrow)}
    assert set(mgot) == set(mwant)
    for key, v in mwant.items():
        assert mgot[key] == pytest.approx(v, rel=1e-14), key
