"""The committed golden fixtures (tests/golden/operator_fixtures.npz, written by
tests/golden/make_golden.py from the oracle's CSR assembly) applied by the
GPU's matrix-free stencil: 2-D and 3-D grids, Neumann top on/off, Sommerfeld
on/off, BC order 1/2, real and complex omega.  Bar: complex128 relative error
<= 1e-13 against the stored H @ x.  Collected after the other GPU files."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import helmholtz_amd as ha  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "operator_fixtures.npz")


def _cases():
    f = np.load(GOLDEN)
    return sorted({k.split("__")[0] for k in f.files})


@pytest.mark.parametrize("key", _cases())
def test_gpu_apply_matches_golden_fixture(key):
    f = np.load(GOLDEN)
    spec = f[key + "__spec"]
    dim = int(spec[0])
    nodes = [int(v) for v in spec[1:1 + dim]]
    neumann, somm, order = bool(spec[4]), bool(spec[5]), int(spec[6])
    print(f"golden case {key}: nodes={nodes} neumann={neumann} somm={somm} bc={order}", flush=True)
    mesh = ha.getRegularMesh(f[key + "__domain"], [n - 1 for n in nodes])
    w = complex(f[key + "__omega"])
    w = w.real if w.imag == 0.0 else w
    Hg = ha.GetHelmholtzOperator(mesh, f[key + "__m"], w, f[key + "__gamma"], neumann, somm, order)
    y = Hg @ f[key + "__x"]
    yr = f[key + "__y"]
    err = np.linalg.norm(y - yr) / np.linalg.norm(yr)
    print(f"  rel err {err:.2e}")
    assert err <= 1e-13
