"""The host half of getABL (helmholtz.jl_amd/csrc/hh_abl.h: the per-axis ramps,
with the reference's 3-D face mask `code`) compiled with g++ and combined per
node as the device kernel does (2-D: amp (p1 + p2 - p1 p2); 3-D:
min(amp, amp (g1 + g2 + g3))), against the oracle's getABL
(src/GetHelmholtz.jl:97-220) for every face mask of small grids.  CPU only;
the device combination itself is checked in tests/test_gpu_operator.py."""
import ctypes as C
import itertools
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import helmholtz_ref as hr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "helmholtz.jl_amd", "csrc", "hh_abl.h")

SRC = r"""
#include <algorithm>
#include "%s"
extern "C" int abl_host(int dim, const int64_t* n, int neumann, const int64_t* pad, double amp,
                        const int32_t* faces, double* out) {
  try {
    auto P = hh::abl_profiles(dim, n, neumann, pad, faces);
    const int64_t n1 = P[0].size(), n2 = P[1].size(), n3 = P[2].size();
    for (int64_t k = 0; k < n3; ++k)
      for (int64_t j = 0; j < n2; ++j)
        for (int64_t i = 0; i < n1; ++i) {
          double v;
          if (dim == 2) {
            const double a = P[0][i], b = P[1][j];
            v = (a + b - a * b) * amp;
          } else {
            const double g = amp * (P[0][i] + P[1][j] + P[2][k]);
            v = g >= amp ? amp : g;
          }
          out[i + n1 * (j + n2 * k)] = v;
        }
  } catch (...) {
    return 1;
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("abl")
    src, so = d / "abl_host.cpp", d / "abl_host.so"
    src.write_text(SRC % HEADER)
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", str(src), "-o", str(so)], check=True)
    lib = C.CDLL(str(so))
    lib.abl_host.argtypes = [C.c_int, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_int64), C.c_double,
                             C.POINTER(C.c_int32), C.c_void_p]
    return lib


def _run(lib, n, neumann, pad, amp, code):
    n3 = (C.c_int64 * 3)(*(list(n) + [1] * (3 - len(n))))
    p3 = (C.c_int64 * 3)(*(list(pad) + [1] * (3 - len(pad))))
    faces = None if code is None else (C.c_int32 * 6)(*([int(v) for v in np.ravel(code)] + [1] * (6 - np.size(code))))
    out = np.zeros(int(np.prod(n)))
    assert lib.abl_host(len(n), n3, int(neumann), p3, amp, faces, out.ctypes.data) == 0
    return out.reshape(tuple(n), order="F")


@pytest.mark.parametrize("n,pad", [((9, 8, 7), (2, 3, 2)), ((12, 10, 9), (3, 2, 4))])
@pytest.mark.parametrize("neumann", [False, True])
def test_3d_every_face_mask(host, n, pad, neumann):
    for bits in itertools.product([0, 1], repeat=6):
        code = np.array(bits, dtype=bool).reshape(3, 2)
        g = _run(host, n, neumann, pad, 2.5, code)
        r = hr.get_abl(n, neumann, pad, 2.5, code)
        np.testing.assert_allclose(g, r, rtol=1e-14, atol=1e-14, err_msg=str(bits))
    # no mask == every face on
    np.testing.assert_array_equal(_run(host, n, neumann, pad, 2.5, None),
                                  _run(host, n, neumann, pad, 2.5, np.ones((3, 2), bool)))


@pytest.mark.parametrize("neumann", [False, True])
def test_2d_ignores_the_mask(host, neumann):
    n, pad = (12, 10), (3, 4)
    r = hr.get_abl(n, neumann, pad, 2.5)
    for bits in itertools.product([0, 1], repeat=4):
        code = np.array(bits, dtype=bool).reshape(2, 2)
        np.testing.assert_allclose(_run(host, n, neumann, pad, 2.5, code), r, rtol=1e-14, atol=1e-14)
