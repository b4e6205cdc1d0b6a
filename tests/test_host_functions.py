"""The package's host-side copies of the reference's small functions against
the oracle and against hand-computed values:

* getMaximalFrequency (src/GetHelmholtz.jl:75-79),
* getShiftedHelmholtzParam (src/Helmholtz.jl:32-34),
* getTopPointSrc / getMidPointSrc (src/getPointSource.jl:63-79),
* getAcousticPointSource (src/getPointSource.jl:105-111) and its batched
  form point_sources.

These are plain NumPy in helmholtz_amd (no library call), so they run here
without a GPU.  The hand-typed values pin the oracle as well: a wrong constant,
axis or index in either copy fails one of them."""
import math

import numpy as np
import pytest

from oracle import helmholtz_ref as hr

import helmholtz_amd as ha

MESHES = [([0.0, 1.0, 0.0, 1.0], [128, 128]), ([0.0, 13.5, 0.0, 4.2], [256, 128]),
          ([0.0, 2.0, 0.0, 1.0], [63, 31]), ([0.0, 1.0, 0.0, 1.3, 0.0, 0.7], [16, 12, 9]),
          ([0.0, 1.0, 0.0, 1.0, 0.0, 1.0], [8, 8, 8])]


def _pair(dom, n):
    return ha.getRegularMesh(dom, n), hr.get_regular_mesh(dom, n)


@pytest.mark.parametrize("dom,n", MESHES)
def test_maximal_frequency_matches_oracle(dom, n):
    Mp, Mo = _pair(dom, n)
    rng = np.random.default_rng(sum(n))
    v = rng.uniform(1.5, 4.5, int(np.prod(np.asarray(n) + 1)))
    m = 1.0 / v**2
    assert ha.getMaximalFrequency(m, Mp) == hr.get_maximal_frequency(m, Mo)


def test_maximal_frequency_known_answer():
    """v_min = 1.5 on h = (0.01, 0.02): w = 0.2 pi v_min / max(h) = 15 pi,
    i.e. a wavelength 2 pi v_min / w = 0.2 = 10 points of the coarser axis."""
    Mp, Mo = _pair([0.0, 1.0, 0.0, 2.0], [100, 100])
    m = np.full(101 * 101, 1.0 / 3.0**2)
    m[17] = 1.0 / 1.5**2
    for w in (ha.getMaximalFrequency(m, Mp), hr.get_maximal_frequency(m, Mo)):
        assert w == pytest.approx(15.0 * math.pi, rel=1e-15)
        assert 2 * math.pi * 1.5 / w == pytest.approx(10 * 0.02, rel=1e-15)


def test_shifted_param_matches_oracle_and_known_answer():
    """gamma + s Re(w): with gamma = (0.1, 0.2), w = 3 + 1i, s = 0.5 the new
    gamma is (1.6, 1.7); m, w and the flags are unchanged."""
    Mp, Mo = _pair([0.0, 1.0, 0.0, 1.0], [1, 0 + 1])
    g = np.array([0.1, 0.2, 0.3, 0.4])
    m = np.array([1.0, 2.0, 3.0, 4.0])
    for w in (3.0 + 1.0j, 2.5):
        pp = ha.getShiftedHelmholtzParam(ha.HelmholtzParam(Mp, g, m, w, True, False), 0.5)
        po = hr.get_shifted_helmholtz_param(hr.HelmholtzParam(Mo, g, m, w, True, False), 0.5)
        np.testing.assert_array_equal(pp.gamma, po.gamma)
        np.testing.assert_allclose(pp.gamma, g + 0.5 * complex(w).real, rtol=0, atol=1e-15)
        assert pp.m is m and pp.omega == w and pp.NeumannOnTop and not pp.Sommerfeld
    pk = ha.getShiftedHelmholtzParam(ha.HelmholtzParam(Mp, g[:2], m[:2], 3.0 + 1.0j, False, True), 0.5)
    np.testing.assert_allclose(pk.gamma, [1.6, 1.7], rtol=0, atol=1e-15)


@pytest.mark.parametrize("dom,n", MESHES)
def test_source_positions_match_oracle(dom, n):
    Mp, Mo = _pair(dom, n)
    assert list(ha.getTopPointSrc(Mp)) == list(hr.get_top_point_src(Mo))
    assert list(ha.getMidPointSrc(Mp)) == list(hr.get_mid_point_src(Mo))


def test_source_positions_known_answers():
    """div(n_cells + 1, 2) per axis, 1-based; 128 cells (129 nodes) give node 64,
    one node before the centre node 65; the top source sits at depth index 1."""
    Mp, Mo = _pair([0.0, 1.0, 0.0, 1.0], [128, 128])
    for top, mid in ((ha.getTopPointSrc(Mp), ha.getMidPointSrc(Mp)),
                     (hr.get_top_point_src(Mo), hr.get_mid_point_src(Mo))):
        assert list(top) == [64, 1]
        assert list(mid) == [64, 64]
    Mp, Mo = _pair([0.0, 1.0, 0.0, 1.0, 0.0, 1.0], [9, 8, 7])
    for top, mid in ((ha.getTopPointSrc(Mp), ha.getMidPointSrc(Mp)),
                     (hr.get_top_point_src(Mo), hr.get_mid_point_src(Mo))):
        assert list(top) == [5, 4, 1]
        assert list(mid) == [5, 4, 4]


@pytest.mark.parametrize("dom,n", MESHES)
def test_acoustic_point_source_matches_oracle(dom, n):
    Mp, Mo = _pair(dom, n)
    for src in (None, hr.get_mid_point_src(Mo)):
        qp, sp = ha.getAcousticPointSource(Mp, np.complex128, src)
        qo, so = hr.get_acoustic_point_source(Mo, src)
        assert list(sp) == list(so)
        assert qp.shape == tuple(int(v) for v in Mo.nodes)
        np.testing.assert_array_equal(qp.ravel(order="F"), qo)


def test_acoustic_point_source_known_answer():
    """257 x 129 nodes on [0, 13.5] x [0, 4.2] (the reference test's mesh): the
    top source is node (128, 1), column-major index 128, value
    1 / (h1^2 + h2^2) with h = (13.5/256, 4.2/128)."""
    Mp, Mo = _pair([0.0, 13.5, 0.0, 4.2], [256, 128])
    val = 1.0 / ((13.5 / 256) ** 2 + (4.2 / 128) ** 2)
    for q in (ha.getAcousticPointSource(Mp, np.complex128)[0].ravel(order="F"), hr.get_acoustic_point_source(Mo)[0]):
        nz = np.flatnonzero(q)
        assert nz.tolist() == [127]
        assert q[127] == pytest.approx(val, rel=1e-15)


def test_point_sources_block_equals_columns_of_single_sources():
    """point_sources(M, subs) column c == getAcousticPointSource(M, T, subs[c])."""
    Mp, Mo = _pair([0.0, 2.0, 0.0, 1.0], [64, 32])
    subs = [(3 + 4 * k, 2) for k in range(8)] + [(65, 33), (1, 1)]
    B = ha.point_sources(Mp, subs)
    assert B.shape == (65 * 33, len(subs)) and B.flags.f_contiguous
    for c, s in enumerate(subs):
        qo, _ = hr.get_acoustic_point_source(Mo, list(s))
        np.testing.assert_array_equal(B[:, c], qo)
