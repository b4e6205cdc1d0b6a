"""Seeded media generators (the benchmark inputs): fixed seeds give fixed
arrays (SHA-256 of the values rounded well above floating-point noise), the
value ranges and the source layout hold, and a different seed gives a
different medium.  CPU only."""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "helmholtz.jl_amd"))

from helmholtz_amd import media  # noqa: E402


def _digest(a, decimals):
    return hashlib.sha256(np.ascontiguousarray(np.round(a, decimals)).tobytes()).hexdigest()[:16]


def test_grf_medium_checksum_and_statistics():
    v = media.grf_velocity_2d((129, 129), seed=1)
    assert v.shape == (129, 129)
    assert _digest(v, 6) == "652ca2b07758bc9d"
    assert abs(v.mean() - 2.0) < 1e-12 and abs(v.std() - 0.3) < 1e-12
    assert _digest(media.grf_velocity_2d((129, 129), seed=2), 6) != "652ca2b07758bc9d"


def test_layered_medium_checksum_and_range():
    v = media.layered_body_velocity_2d((1025, 513), seed=2)
    assert v.shape == (1025, 513)
    assert _digest(v, 3) == "c53c2648fb20618e"
    assert v.min() >= 1500.0 and v.max() <= 4500.0
    # velocity increases with depth on average (layers from vmin to vmax)
    assert v[:, :50].mean() < v[:, -50:].mean()
    v2 = media.layered_body_velocity_2d((1025, 513), seed=2)
    assert np.array_equal(v, v2)


def test_extruded_3d_medium_checksum_and_range():
    v = media.layered_body_velocity_3d((65, 49, 33), seed=3)
    assert v.shape == (65, 49, 33)
    assert _digest(v, 3) == "e4eae054a5e5c908"
    assert v.min() >= 1500.0 and v.max() <= 4500.0


def test_frequency_and_sources():
    w = media.omega_for_ppw(1500.0, 10.0, 8)
    assert abs(w - 2 * np.pi * 1500.0 / 80.0) < 1e-12  # 18.75 Hz
    assert abs(2 * np.pi * 1500.0 / w / 10.0 - 8.0) < 1e-12
    src = media.surface_sources_2d((1025, 513), 256)
    assert src[0] == (3, 2) and src[-1] == (3 + 4 * 255, 2) and len(src) == 256
