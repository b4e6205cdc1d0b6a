"""Writes tests/golden/operator_fixtures.npz: small seeded inputs (m, gamma,
omega, x) and the oracle's H @ x for a spread of operator options (2-D/3-D,
Neumann top on/off, Sommerfeld on/off, BC order 1/2, complex omega).

The reference holds no golden vectors of its own; these fixtures freeze the
oracle's output so a later change to it is caught bit for bit
(tests/test_oracle.py) and give the GPU tests inputs that travel without the
oracle's assembly step.  Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import helmholtz_ref as hr  # noqa: E402

CASES = {
    "c2d_a": ([9, 7], False, True, 2, False),
    "c2d_b": ([9, 7], True, True, 1, True),
    "c2d_c": ([16, 11], True, False, 2, False),
    "c3d_a": ([6, 5, 4], False, True, 2, False),
    "c3d_b": ([6, 5, 4], True, True, 1, True),
}


def main():
    out = {}
    rng = np.random.default_rng(2024)
    for key, (nodes, neumann, somm, order, cplx) in CASES.items():
        dim = len(nodes)
        dom = np.array(sum(([0.0, 1.0 + 0.5 * d] for d in range(dim)), []))
        M = hr.get_regular_mesh(dom, [n - 1 for n in nodes])
        N = int(np.prod(nodes))
        m = rng.uniform(0.05, 0.45, N)
        w = hr.get_maximal_frequency(m, M) * (complex(1, -0.2) if cplx else 1.0)
        g = 0.01 * abs(w) + hr.get_abl(nodes, neumann, [2] * dim, abs(w)).ravel(order="F")
        x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        H = hr.get_helmholtz_operator7(M, m, w, g, neumann, somm, order)
        spec = np.array([dim] + nodes + [0] * (3 - dim) + [int(neumann), int(somm), order])
        out.update({key + "__spec": spec, key + "__domain": dom, key + "__m": m, key + "__gamma": g,
                    key + "__omega": np.complex128(w), key + "__x": x, key + "__y": H @ x})
    np.savez_compressed(os.path.join(HERE, "operator_fixtures.npz"), **out)


if __name__ == "__main__":
    main()
