"""Shared problem builders for the tests: the same inputs go to the oracle
(SciPy CSR / textbook SL-MG) and to the HIP path."""
from __future__ import annotations

import numpy as np

from oracle import helmholtz_ref as hr


def random_medium(nodes, seed, vmin=1.5, vmax=4.5):
    rng = np.random.default_rng(seed)
    v = rng.uniform(vmin, vmax, size=tuple(nodes))
    return (1.0 / v**2).ravel(order="F")


def small_case(nodes, seed=0, dim=None, neumann=True, sommerfeld=True, order=2, omega=None, pad=None):
    """(mesh, m, gamma, omega) for a random medium on a small grid."""
    nodes = list(nodes)
    dim = len(nodes)
    dom = []
    for d in range(dim):
        dom += [0.0, 1.0 + 0.37 * d]
    M = hr.get_regular_mesh(dom, [n - 1 for n in nodes])
    m = random_medium(nodes, seed)
    w = hr.get_maximal_frequency(m, M) if omega is None else omega
    pad = [max(1, n // 6) for n in nodes] if pad is None else pad
    g0 = 0.01 * abs(complex(w)) * np.ones(int(np.prod(nodes)))
    abl = hr.get_abl(nodes, neumann, pad, abs(complex(w)).real).ravel(order="F")
    return M, m, g0 + abl, w
