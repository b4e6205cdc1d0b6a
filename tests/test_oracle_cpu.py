"""The oracle's C/OpenMP restatement (oracle/slmg_cpu.c) against the NumPy
oracle: the assembled operator to 1e-13 and the plain solve with the same
iteration counts and iterates to 1e-9 (V, W and the hybrid W/V cycle, 2-D and
3-D, both boundary orders).  CPU only."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle import build_cpu, helmholtz_ref as hr, slmg_ref as sr  # noqa: E402
from tests.problems import small_case  # noqa: E402


@pytest.fixture(scope="module")
def cpu():
    if not os.path.exists(os.path.join(os.path.dirname(build_cpu.__file__), "libslmg_cpu.so")):
        build_cpu.build()
    from oracle import slmg_cpu

    return slmg_cpu


@pytest.mark.parametrize("nodes,neumann,sommerfeld,order,omega", [
    ((17, 9), True, True, 2, None),
    ((17, 9), False, True, 1, None),
    ((13, 11), True, False, 2, 30.0 + 2.0j),
    ((9, 7, 5), True, True, 2, None),
    ((9, 7, 5), False, False, 1, None),
])
def test_assembled_operator_matches_kron(cpu, nodes, neumann, sommerfeld, order, omega):
    M, m, g, w = small_case(nodes, seed=3, neumann=neumann, omega=omega)
    p = hr.HelmholtzParam(M, g, m, w, neumann, sommerfeld)
    H = hr.operator_from_param(p, order)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(m.size) + 1j * rng.standard_normal(m.size)
    y = cpu.apply(p, x, order)
    ref = H @ x
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) <= 1e-13


@pytest.mark.parametrize("nodes,levels,cycle,wlevels", [
    ((65, 33), 3, "V", 0),
    ((65, 33), 4, "W", 0),
    ((65, 33), 4, "W", 2),
    ((17, 17, 9), 2, "V", 0),
])
def test_plain_solve_matches_numpy_oracle(cpu, nodes, levels, cycle, wlevels):
    M, m, g, w = small_case(nodes, seed=7)
    p = hr.HelmholtzParam(M, g, m, w, True, True)
    n1 = int(M.nodes[0])
    b = np.zeros(m.size, dtype=complex)
    b[n1 // 2 + n1] = 1.0 / M.h[0] ** 2
    hier = sr.build_hierarchy(p, levels, 0.5)
    H = hr.operator_from_param(p)
    Mf = lambda v: sr.vcycle(hier, 0, v, 0.8, 2, 2, cycle, wlevels)
    x_ref, it_ref, _ = sr.fgmres(lambda v: H @ v, b, Mf, 10, 40, 1e-8)
    x, its, _, _ = cpu.solve(p, b, levels=levels, shift=0.5, inner=10, max_outer=40, tol=1e-8,
                             cycle=cycle, wlevels=wlevels)
    assert its[0] == it_ref
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) <= 1e-9
    assert np.linalg.norm(H @ x - b) / np.linalg.norm(b) <= 1e-8


def test_multi_rhs_and_zero_rhs(cpu):
    M, m, g, w = small_case((33, 17), seed=2)
    p = hr.HelmholtzParam(M, g, m, w, True, True)
    B = np.zeros((m.size, 3), dtype=complex)
    B[40, 0] = 1.0
    B[100, 2] = 2.0 - 1.0j
    X, its, _, _ = cpu.solve(p, B, levels=2, tol=1e-7)
    H = hr.operator_from_param(p)
    assert its[1] == 0 and np.all(X[:, 1] == 0)
    for c in (0, 2):
        assert np.linalg.norm(H @ X[:, c] - B[:, c]) / np.linalg.norm(B[:, c]) <= 1e-7
