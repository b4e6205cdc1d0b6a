"""GPU parity of the matrix-free stencil (hh_op_apply) against the oracle's
assembled CSR operator (the reference's kron recipe, src/GetHelmholtz.jl:33-50).

Bar: complex128 relative error <= 1e-13 (the parity bar for the stencil
apply); complex64 <= 2e-6.  Relative error is ||y_gpu - y_ref|| / ||y_ref||.
"""
import numpy as np
import pytest

from oracle import helmholtz_ref as hr
from problems import small_case

pytestmark = pytest.mark.gpu

import helmholtz_amd as ha  # noqa: E402


def _hh_mesh(M):
    return ha.getRegularMesh(M.domain, M.n)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


CASES = [
    # nodes, neumann, sommerfeld, order, complex omega
    ((17, 13), False, True, 2, False),
    ((17, 13), True, True, 2, False),
    ((33, 9), True, False, 2, False),
    ((12, 10), False, False, 1, False),
    ((12, 10), True, True, 1, True),
    ((129, 129), False, True, 2, False),
    ((1025, 513), True, True, 2, False),
    ((9, 8, 7), False, True, 2, False),
    ((9, 8, 7), True, True, 1, True),
    ((33, 17, 21), True, True, 2, False),
    ((65, 65, 33), True, False, 2, False),
]


@pytest.mark.parametrize("nodes,neumann,sommerfeld,order,cplx", CASES)
@pytest.mark.parametrize("shift", [0.0, 0.5])
def test_apply_c128_matches_csr(nodes, neumann, sommerfeld, order, cplx, shift):
    print(f"case nodes={nodes} neumann={neumann} somm={sommerfeld} bc={order} cplx={cplx} shift={shift}", flush=True)
    M, m, g, w = small_case(nodes, seed=sum(nodes), neumann=neumann)
    if cplx:
        w = complex(w, -0.3 * w)
    if shift and cplx:
        pytest.skip("GetHelmholtzShiftOP is defined for real omega only")
    H = hr.get_helmholtz_operator7(M, m, w, g, neumann, sommerfeld, order)
    if shift:
        H = H + hr.get_helmholtz_shift_op(m, w, shift)
    Hg = ha.GetHelmholtzOperator(_hh_mesh(M), m, w, g, neumann, sommerfeld, order)
    if shift:
        Hg = Hg + ha.GetHelmholtzShiftOP(m, w, shift)
    rng = np.random.default_rng(7)
    N = H.shape[0]
    X = np.asfortranarray(rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3)))
    Y = Hg @ X
    Yr = H @ X
    for c in range(3):
        assert _rel(Y[:, c], Yr[:, c]) <= 1e-13


@pytest.mark.parametrize("nodes", [(129, 129), (9, 8, 7), (65, 33, 17)])
def test_apply_c64(nodes):
    M, m, g, w = small_case(nodes, seed=3, neumann=True)
    H = hr.get_helmholtz_operator7(M, m, w, g, True, True, 2)
    Hg = ha.GetHelmholtzOperator(_hh_mesh(M), m, w, g, True, True, 2)
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(H.shape[0]) + 1j * rng.standard_normal(H.shape[0])).astype(np.complex64)
    y = Hg.apply(x, prec="c64")
    assert y.dtype == np.complex64
    assert _rel(y.astype(np.complex128), H @ x.astype(np.complex128)) <= 2e-6


def test_apply_device_pointers():
    from helmholtz_amd.device import DeviceArray
    M, m, g, w = small_case((257, 129), seed=5, neumann=True)
    H = hr.get_helmholtz_operator7(M, m, w, g, True, True, 2)
    Hg = ha.GetHelmholtzOperator(_hh_mesh(M), m, w, g, True, True, 2)
    rng = np.random.default_rng(2)
    N = H.shape[0]
    X = np.asfortranarray(rng.standard_normal((N, 4)) + 1j * rng.standard_normal((N, 4)))
    Y = (Hg @ DeviceArray.from_numpy(X)).to_numpy()
    for c in range(4):
        assert _rel(Y[:, c], H @ X[:, c]) <= 1e-13


@pytest.mark.parametrize("n,neumann,pad", [((12, 10), False, (3, 4)), ((12, 10), True, (3, 4)),
                                           ((40, 30), True, (20, 15)), ((9, 8, 7), False, (2, 3, 2)),
                                           ((9, 8, 7), True, (3, 3, 3))])
def test_abl_matches_oracle(n, neumann, pad):
    g = ha.getABL(n, neumann, pad, 2.5)
    r = hr.get_abl(n, neumann, pad, 2.5)
    assert g.shape == r.shape
    np.testing.assert_allclose(g, r, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("nodes,neumann,order", [((12, 10), False, 2), ((12, 10), True, 2), ((33, 17), True, 1),
                                                  ((9, 8, 7), False, 2), ((9, 8, 7), True, 1)])
def test_sommerfeld_bc_matches_oracle(nodes, neumann, order):
    """getSommerfeldBC (src/GetHelmholtz.jl:222-247) on the GPU against the
    oracle's face-by-face sums: zero in the interior, -i w BC/h sqrt(m) on
    faces, sums at edges and corners, no top face with a Neumann top."""
    M, m, g, w = small_case(nodes, seed=12, neumann=neumann)
    r = hr.get_sommerfeld_bc(M, m, w, neumann, order)
    s = ha.getSommerfeldBC(_hh_mesh(M), m, w, neumann, order)
    assert s.shape == r.shape
    np.testing.assert_allclose(s, r, rtol=1e-15, atol=0)
    assert np.count_nonzero(s) == np.count_nonzero(r)
    with pytest.raises(TypeError):
        ha.getSommerfeldBC(_hh_mesh(M), m, complex(w, 0.1), neumann, order)


def test_eight_argument_form_returns_gamma():
    M, m, g, w = small_case((33, 17), seed=9, neumann=True)
    g0 = 0.01 * w * np.ones(m.size)
    Hr, gr = hr.get_helmholtz_operator(M, m, w, g0, True, [5, 4], w, True)
    Hg, gg = ha.GetHelmholtzOperator(_hh_mesh(M), m, w, g0, True, [5, 4], w, True)
    np.testing.assert_allclose(gg, gr, rtol=1e-14, atol=1e-14)
    x = np.random.default_rng(0).standard_normal(m.size) + 0j
    assert _rel(Hg @ x, Hr @ x) <= 1e-13


@pytest.mark.parametrize("nodes,beta,neumann,sommerfeld,shift", [
    ((33, 17), 2.0 / 3.0, True, True, 0.0),
    ((33, 17), 0.3, False, True, 0.5),
    ((20, 13), 1.0, True, False, 0.0),
    ((9, 8, 7), (0.5, 0.25), True, True, 0.0),
    ((9, 8, 7), (1.0, 1.0), False, True, 0.5),
])
def test_ho_apply_matches_csr(nodes, beta, neumann, sommerfeld, shift):
    """GetHelmholtzOperatorHO (src/GetHelmholtz.jl:54-72) matrix-free on the
    GPU against the oracle's kron-assembled CSR, complex128, 1e-13."""
    M, m, g, w = small_case(nodes, seed=21, neumann=neumann)
    H = hr.get_helmholtz_operator_ho(M, m, w, g, neumann, sommerfeld, beta)
    if shift:
        H = H + hr.get_helmholtz_shift_op(m, w, shift)
    mesh = ha.getRegularMesh(M.domain, M.n)
    Hg = ha.GetHelmholtzOperatorHO(mesh, m, w, g, neumann, sommerfeld, beta)
    if shift:
        Hg = Hg + ha.GetHelmholtzShiftOP(m, w, shift)
    rng = np.random.default_rng(22)
    X = np.asfortranarray(rng.standard_normal((m.size, 3)) + 1j * rng.standard_normal((m.size, 3)))
    Y = Hg @ X
    ref = H @ X
    err = np.linalg.norm(Y - ref) / np.linalg.norm(ref)
    assert err <= 1e-13, err
