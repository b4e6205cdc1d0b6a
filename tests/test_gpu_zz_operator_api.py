"""getHelmholtzFun (src/GetHelmholtz.jl:85-95) on the GPU: the outer operator
built from the shifted operator and the negated shift, SH + (-i s w^2 m) = H,
applied into a caller's buffer, against the oracle's CSR H (1e-13).
Collected after the other GPU files."""
import numpy as np
import pytest

from oracle import helmholtz_ref as hr
from problems import small_case

pytestmark = pytest.mark.gpu

import helmholtz_amd as ha  # noqa: E402


@pytest.mark.parametrize("nodes", [(33, 17), (9, 8, 7)])
def test_helmholtz_fun_applies_h(nodes):
    print(f"getHelmholtzFun {nodes}", flush=True)
    M, m, g, w = small_case(nodes, seed=31, neumann=True)
    H = hr.get_helmholtz_operator7(M, m, w, g, True, True)
    Hg = ha.GetHelmholtzOperator(ha.getRegularMesh(M.domain, M.n), m, w, g, True, True)
    SH = Hg + ha.GetHelmholtzShiftOP(m, w, 0.5)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(m.size) + 1j * rng.standard_normal(m.size)
    y = np.zeros(m.size, complex)
    Hfun = ha.getHelmholtzFun(SH, ha.GetHelmholtzShiftOP(m, w, -0.5), y, 1)
    out = Hfun(x)
    assert out is y
    ref = H @ x
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) <= 1e-13
    # a block of right-hand sides into an N x 2 buffer
    X = np.asfortranarray(rng.standard_normal((m.size, 2)) + 1j * rng.standard_normal((m.size, 2)))
    Y = np.zeros((m.size, 2), complex, order="F")
    ha.getHelmholtzFun(SH, ha.GetHelmholtzShiftOP(m, w, -0.5), Y, 1)(X)
    for c in range(2):
        r = H @ X[:, c]
        assert np.linalg.norm(Y[:, c] - r) / np.linalg.norm(r) <= 1e-13
