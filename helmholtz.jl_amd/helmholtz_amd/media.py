"""Seeded synthetic media and workloads for the Helmholtz benchmarks.

These generators have no counterpart in Helmholtz.jl (its tests build constant
media inline, e.g. test/ShiftedLaplacianTest.jl:15); they define the benchmark
inputs once, deterministically from a seed.  Arrays are nodal and column-major:
shape ``(n1, n2[, n3])`` with axis 0 = x (fastest), the last axis = depth, and
depth index 0 is the "top" face (src/GetHelmholtz.jl:151-155, 230-231).
"""
from __future__ import annotations

import math

import numpy as np


def gaussian_random_field(shape, h, corr_len, seed):
    """Unit-std stationary Gaussian random field: white noise filtered in
    Fourier space by exp(-(k corr_len)^2 / 4) (a Gaussian covariance with
    correlation length ``corr_len``), then normalised to zero mean, unit std."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    F = np.fft.fftn(noise)
    ks = np.meshgrid(*[2 * np.pi * np.fft.fftfreq(n, d=hh) for n, hh in zip(shape, h)],
                     indexing="ij")
    k2 = sum(k * k for k in ks)
    field = np.real(np.fft.ifftn(F * np.exp(-k2 * corr_len**2 / 4.0)))
    field -= field.mean()
    return field / field.std()


def grf_velocity_2d(n_nodes=(129, 129), seed=1, v0=2.0, amp=0.3, corr_len=0.1, domain=(0, 1, 0, 1)):
    """v = v0 + amp * G on the unit square (the small plain-method workload)."""
    h = [(domain[1] - domain[0]) / (n_nodes[0] - 1), (domain[3] - domain[2]) / (n_nodes[1] - 1)]
    return v0 + amp * gaussian_random_field(tuple(n_nodes), h, corr_len, seed)


def layered_body_velocity_2d(n_nodes=(1025, 513), seed=2, vmin=1500.0, vmax=4500.0,
                             n_interfaces=12, noise=0.05):
    """Layered medium from ``n_interfaces`` seeded random interfaces with
    velocities increasing (with jitter) from vmin to vmax, a seeded elliptic
    high-velocity body at vmax, and multiplicative Gaussian noise; clipped to
    [vmin, vmax] so the points-per-wavelength target holds everywhere."""
    rng = np.random.default_rng(seed)
    n1, n2 = n_nodes
    x = np.linspace(0.0, 1.0, n1)
    z = np.linspace(0.0, 1.0, n2)
    depths = np.sort(rng.uniform(0.06, 0.95, n_interfaces))
    layer_v = np.linspace(vmin, vmax, n_interfaces + 1)
    layer_v[1:-1] += rng.uniform(-0.08, 0.08, n_interfaces - 1) * (vmax - vmin) / n_interfaces
    v = np.full((n1, n2), layer_v[0])
    Z = z[None, :]
    for k in range(n_interfaces):
        # smooth undulating interface: a few random sinusoids
        amp = rng.uniform(0.005, 0.03, 3)
        freq = rng.uniform(1.0, 6.0, 3)
        phase = rng.uniform(0.0, 2 * np.pi, 3)
        zk = depths[k] + sum(a * np.sin(2 * np.pi * f * x + p) for a, f, p in zip(amp, freq, phase))
        v = np.where(Z >= zk[:, None], layer_v[k + 1], v)
    # elliptic body
    cx, cz = rng.uniform(0.3, 0.7), rng.uniform(0.4, 0.7)
    ax, az = rng.uniform(0.08, 0.18), rng.uniform(0.05, 0.12)
    th = rng.uniform(0.0, np.pi)
    X = x[:, None]
    xr = (X - cx) * np.cos(th) + (Z - cz) * np.sin(th)
    zr = -(X - cx) * np.sin(th) + (Z - cz) * np.cos(th)
    v = np.where((xr / ax) ** 2 + (zr / az) ** 2 <= 1.0, vmax, v)
    v = v * (1.0 + noise * rng.standard_normal(v.shape))
    return np.clip(v, vmin, vmax)


def layered_body_velocity_3d(n_nodes=(193, 193, 97), seed=3, vmin=1500.0, vmax=4500.0):
    """3-D medium: the 2-D layered+body generator on the (x, depth) plane,
    extruded along y with seeded smooth lateral shifts of the x coordinate."""
    n1, n2, n3 = n_nodes
    rng = np.random.default_rng(seed)
    base = layered_body_velocity_2d((2 * n1, n3), seed=seed * 7919 + 1, vmin=vmin, vmax=vmax,
                                    noise=0.0)
    y = np.linspace(0.0, 1.0, n2)
    shift = (0.1 * np.sin(2 * np.pi * rng.uniform(0.5, 1.5) * y + rng.uniform(0, 2 * np.pi))
             + 0.05 * np.sin(2 * np.pi * rng.uniform(1.5, 3.0) * y + rng.uniform(0, 2 * np.pi)))
    v = np.empty((n1, n2, n3))
    ix = np.arange(n1)
    for j in range(n2):
        src = np.clip(np.round(ix + (0.25 + shift[j]) * n1).astype(int), 0, 2 * n1 - 1)
        v[:, j, :] = base[src, :]
    v = v * (1.0 + 0.05 * rng.standard_normal(v.shape))
    return np.clip(v, vmin, vmax)


def omega_for_ppw(vmin, h, ppw):
    """Angular frequency giving ``ppw`` points per wavelength at ``vmin``."""
    return 2.0 * math.pi * vmin / (ppw * h)


def surface_sources_2d(n_nodes, count, depth_index=2, first=3, step=4):
    """1-based (ix, iz) subscripts of ``count`` sources on depth row ``depth_index``."""
    return [(first + step * k, depth_index) for k in range(count)]
