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


def attenuation_case(nn=129):
    """The flow of test/AttenuationTest.jl at 129x129: unit slowness on
    [0, 10]^2, 20 points per wavelength, a point source 1/h^2 at node
    div(n, 2) on both axes (1-based, as the reference test's src), no
    Sommerfeld, no Neumann top; attenuation alpha = 0.2 w either added to gamma
    at real w (s1) or carried by the complex frequency w - i alpha/2 (s2)."""
    n = (nn, nn)
    M = hr.get_regular_mesh([0.0, 10.0, 0.0, 10.0], [nn - 1, nn - 1])
    m = np.ones(nn * nn)
    wmax = hr.get_maximal_frequency(m, M)
    w = wmax / 2.0
    alpha = 0.2 * w
    pad = nn * 35 // 512 + 1
    gam = hr.get_abl(n, False, [pad, pad], wmax).ravel(order="F")
    src = [nn // 2, nn // 2]
    q = np.zeros(nn * nn, complex)
    q[hr.loc2cs(n, src) - 1] = 1.0 / M.h[0] ** 2
    return M, m, w, alpha, gam, pad, src, q


def attenuation_decay_slope(s, nn, src, pad, h):
    """Slope of log(|s| sqrt(r)) along the source row between the near field
    and the absorbing layer: the 2-D Green's function decays as
    r^(-1/2) exp(-Im(kappa) r)."""
    S = np.asarray(s).reshape((nn, nn), order="F")
    rs = np.arange(5, nn // 2 - pad - 2)
    r = rs * h
    a = np.abs(S[src[0] - 1 + rs, src[1] - 1])
    return np.polyfit(r, np.log(a * np.sqrt(r)), 1)[0]



def seg_salt_case():
    """The medium of the commented-out line of test/ShiftedLaplacianTest.jl:14
    (examples/SEGmodel2Dsalt.dat, committed as tests/golden/seg_salt_2d.npz)
    on that test's set-up: velocity in km/s transposed to x-fastest nodal
    order (256 x 128), extended by one node on the far x and bottom edges to
    the test's own 257 x 129 grid (even cell counts, so it coarsens) on
    [0, 13.5] x [0, 4.2]; f = 2.5; gamma = 0.01 w plus an absorbing layer of 16
    cells with amplitude getMaximalFrequency; Neumann top, Sommerfeld; the
    test's source (div(n, 2), 1) scaled by 1/h1^2.
    Returns (mesh, m, gamma, w, H, q)."""
    import hashlib
    import os

    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "seg_salt_2d.npz"))
    v = f["v"].astype(np.float64)
    assert hashlib.sha256(v.tobytes()).hexdigest() == str(f["sha256"])
    v = np.pad(v.T * 1e-3, ((0, 1), (0, 1)), mode="edge")
    n = v.shape
    M = hr.get_regular_mesh([0.0, 13.5, 0.0, 4.2], [n[0] - 1, n[1] - 1])
    m = (1.0 / v**2).ravel(order="F")
    w = 2 * np.pi * 2.5
    H, g = hr.get_helmholtz_operator(M, m, w, 0.01 * w * np.ones(m.size), True, [16, 16],
                                     hr.get_maximal_frequency(m, M), True)
    q = np.zeros(m.size, complex)
    q[hr.loc2cs(n, [n[0] // 2, 1]) - 1] = 1.0 / M.h[0] ** 2
    return M, m, g.ravel(order="F"), w, H, q


def c2_case(nsrc=256):
    """The batched 2-D workload: 1025 x 513 nodes, h = 10 m, seeded layered
    medium with an elliptic body (seed 2), 8 points per wavelength at
    1500 m/s, attenuation 0.01 w plus a 20-cell absorbing layer of amplitude
    w, Neumann top, Sommerfeld sides; nsrc point sources on depth row 2
    (x = 3 + 4k).  Returns (oracle mesh, m, gamma, w, CSR H, B)."""
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "helmholtz.jl_amd"))
    from helmholtz_amd import media
    n, h = (1025, 513), 10.0
    v = media.layered_body_velocity_2d(n, seed=2)
    m = (1.0 / v**2).ravel(order="F")
    M = hr.get_regular_mesh([0.0, h * (n[0] - 1), 0.0, h * (n[1] - 1)], [n[0] - 1, n[1] - 1])
    w = media.omega_for_ppw(1500.0, h, 8)
    H, g = hr.get_helmholtz_operator(M, m, w, 0.01 * w * np.ones(m.size), True, [20, 20], w, True)
    B = np.zeros((m.size, nsrc), dtype=complex, order="F")
    for k, (ix, iz) in enumerate(media.surface_sources_2d(n, nsrc)):
        B[hr.loc2cs(M.nodes, [ix, iz]) - 1, k] = 1.0 / np.linalg.norm(M.h) ** 2
    return M, m, g.ravel(order="F"), w, H.tocsr(), B
