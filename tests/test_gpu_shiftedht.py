"""The ShiftedHT argument of solveLinearSystem selects the operator, as the
reference's first solve builds its hierarchy from it
(src/ShiftedLaplacianMultigridSolver.jl:65): an orderNeumannBC = 1 operator is
solved as such, the reference's own sparse(SH') is accepted after a check
against helmParam, and operators of another medium are refused."""
import numpy as np
import pytest

from oracle import helmholtz_ref as hr
from problems import small_case

pytestmark = pytest.mark.gpu

import helmholtz_amd as ha  # noqa: E402


def _setup(seed=61, nodes=(65, 33)):
    M, m, g, w = small_case(nodes, seed=seed, neumann=True)
    mesh = ha.getRegularMesh(M.domain, M.n)
    hp = ha.HelmholtzParam(mesh, g, m, w, True, True)
    mg = ha.getMGparam(complex, int, 3, 1, 100, 1e-6, "Jac", 0.8, 2, 2, "V", "NoMUMPS", 0.5, 0.0)
    s = ha.getShiftedLaplacianMultigridSolver(hp, mg, 0.5, "GMRES", 10)
    return M, m, g, w, mesh, hp, s


def test_bc1_operator_is_solved():
    """GetHelmholtzOperator(helmParam, 1) passed as ShiftedHT: the true
    residual of the orderNeumannBC = 1 system (oracle CSR) meets the
    tolerance, and the BC = 2 system's residual of that solution does not."""
    M, m, g, w, mesh, hp, s = _setup()
    H1 = hr.get_helmholtz_operator7(M, m, w, g, True, True, 1)
    H2 = hr.get_helmholtz_operator7(M, m, w, g, True, True, 2)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(m.size) + 1j * rng.standard_normal(m.size)
    op1 = ha.GetHelmholtzOperator(hp, 1)
    x, s = ha.solveLinearSystem(op1, b, s)
    r1 = np.linalg.norm(H1 @ x - b) / np.linalg.norm(b)
    r2 = np.linalg.norm(H2 @ x - b) / np.linalg.norm(b)
    print(f"BC=1 ShiftedHT: residual of the BC=1 system {r1:.2e}, of the BC=2 system {r2:.2e}")
    assert r1 <= 1e-6
    assert r2 > 1e-3
    # the shifted form the reference's tests pass (H + GetHelmholtzShiftOP(m, w, shift[1])) is accepted too
    sh = op1 + ha.GetHelmholtzShiftOP(m, w, 0.5)
    x2, s = ha.solveLinearSystem(sh, b, s)
    assert np.linalg.norm(H1 @ x2 - b) / np.linalg.norm(b) <= 1e-6


def test_sparse_shifted_adjoint_accepted_and_checked():
    """The reference's sparse(SH') (SH = H + shift op) is matched to helmParam
    (BC 2 and BC 1 alike); a matrix of another medium is refused."""
    M, m, g, w, mesh, hp, s = _setup(seed=62)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(m.size) + 1j * rng.standard_normal(m.size)
    for order in (2, 1):
        H = hr.get_helmholtz_operator7(M, m, w, g, True, True, order)
        SHt = (H + hr.get_helmholtz_shift_op(m, w, 0.5)).conj().T.tocsc()
        s2 = ha.copySolver(s)
        x, s2 = ha.solveLinearSystem(SHt, b, s2)
        res = np.linalg.norm(H @ x - b) / np.linalg.norm(b)
        print(f"sparse SH' (BC={order}): true residual {res:.2e}")
        assert res <= 1e-6
    m_other = m * 1.1
    Ho = hr.get_helmholtz_operator7(M, m_other, w, g, True, True, 2)
    SHo = (Ho + hr.get_helmholtz_shift_op(m_other, w, 0.5)).conj().T.tocsc()
    with pytest.raises(ValueError):
        ha.solveLinearSystem(SHo, b, ha.copySolver(s))


def test_operator_of_other_fields_refused():
    """A matrix-free operator built from another slowness, frequency or
    boundary flag than helmParam's, or with a foreign shift, is refused."""
    M, m, g, w, mesh, hp, s = _setup(seed=63)
    b = np.ones(m.size, complex)
    bad = [ha.GetHelmholtzOperator(mesh, m * 1.05, w, g, True, True),
           ha.GetHelmholtzOperator(mesh, m, 1.01 * w, g, True, True),
           ha.GetHelmholtzOperator(mesh, m, w, g, False, True),
           ha.GetHelmholtzOperator(mesh, m, w, g, True, True) + ha.GetHelmholtzShiftOP(m, w, 0.3)]
    for op in bad:
        with pytest.raises(ValueError):
            ha.solveLinearSystem(op, b, ha.copySolver(s))
    x, _ = ha.solveLinearSystem(ha.GetHelmholtzOperator(hp), b, s)
    H = hr.get_helmholtz_operator7(M, m, w, g, True, True, 2)
    assert np.linalg.norm(H @ x - b) / np.linalg.norm(b) <= 1e-6
