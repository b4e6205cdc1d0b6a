"""The allocation-free operator apply (hh_op_apply_noalloc) against
hh_op_apply, called through the C ABI: bit-identical results on two calls in
a row, the second with a smaller input, in complex128 and complex64, with and
without a shift; a block above the scratch limit is refused before anything
runs on the GPU."""
import ctypes as C

import numpy as np
import pytest

from problems import small_case

pytestmark = pytest.mark.gpu

import helmholtz_amd as ha  # noqa: E402
from helmholtz_amd import _lib as L  # noqa: E402

LIMIT = 32 << 20  # bytes per staged block


def _apply(fn, op, shift, X, prec):
    Y = np.empty_like(X, order="F")
    nrhs = X.shape[1]
    if fn == "noalloc":
        rc = L.lib().hh_op_apply_noalloc(op._h.ptr, shift, X.ctypes.data, Y.ctypes.data, nrhs, prec)
    else:
        rc = L.lib().hh_op_apply(op._h.ptr, shift, X.ctypes.data, Y.ctypes.data, nrhs, prec, L.HH_HOST)
    L.check(rc)
    return Y


@pytest.mark.parametrize("prec", [L.HH_C128, L.HH_C64])
def test_noalloc_apply_equals_apply(prec):
    M, m, g, w = small_case((257, 129), seed=71, neumann=True)
    mesh = ha.getRegularMesh(M.domain, M.n)
    op = ha.GetHelmholtzOperator(mesh, m, w, g, True, True)
    dt = np.complex128 if prec == L.HH_C128 else np.complex64
    rng = np.random.default_rng(5)
    N = m.size
    for nrhs, shift in ((6, 0.0), (2, 0.5)):  # the second call is smaller
        X = np.asfortranarray((rng.standard_normal((N, nrhs)) + 1j * rng.standard_normal((N, nrhs))).astype(dt))
        Y0 = _apply("apply", op, shift, X, prec)
        Y1 = _apply("noalloc", op, shift, X, prec)
        assert np.array_equal(Y0, Y1)
        assert np.linalg.norm(Y1) > 0


def test_noalloc_apply_refuses_blocks_above_the_limit():
    M, m, g, w = small_case((257, 129), seed=72, neumann=True)
    mesh = ha.getRegularMesh(M.domain, M.n)
    op = ha.GetHelmholtzOperator(mesh, m, w, g, True, True)
    N = m.size
    nrhs = LIMIT // (16 * N) + 1  # one column more than fits in complex128
    X = np.zeros((N, nrhs), dtype=np.complex128, order="F")
    Y = np.empty_like(X)
    rc = L.lib().hh_op_apply_noalloc(op._h.ptr, 0.0, X.ctypes.data, Y.ctypes.data, nrhs, L.HH_C128)
    assert rc == L.HH_ERR_ARG
    assert b"exceeds" in L.lib().hh_last_error()
    # the largest block that fits is accepted and equals hh_op_apply
    nfit = LIMIT // (16 * N)
    X = np.asfortranarray(np.random.default_rng(6).standard_normal((N, nfit)).astype(np.complex128))
    assert np.array_equal(_apply("noalloc", op, 0.0, X, L.HH_C128), _apply("apply", op, 0.0, X, L.HH_C128))
