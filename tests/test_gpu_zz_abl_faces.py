"""getABL with the reference's face mask `code` (src/GetHelmholtz.jl:97,
164-218) on the GPU (hh_abl_faces) against the oracle: in 3-D each face of
each axis can be switched off; the 2-D branch does not read the mask.
Collected after the other GPU files."""
import numpy as np
import pytest

from oracle import helmholtz_ref as hr

pytestmark = pytest.mark.gpu

import helmholtz_amd as ha  # noqa: E402

MASKS = [
    [[1, 1], [1, 1], [1, 1]],
    [[0, 1], [1, 1], [1, 1]],
    [[1, 0], [0, 1], [1, 0]],
    [[0, 0], [1, 1], [0, 1]],
    [[0, 0], [0, 0], [0, 0]],
]


@pytest.mark.parametrize("neumann", [False, True])
@pytest.mark.parametrize("mask", range(len(MASKS)))
def test_abl_face_mask_3d(mask, neumann):
    code = np.array(MASKS[mask], dtype=bool)
    n, pad = (9, 8, 7), (2, 3, 2)
    print(f"abl faces 3-D mask {MASKS[mask]} neumann={neumann}", flush=True)
    g = ha.getABL(n, neumann, pad, 2.5, code)
    r = hr.get_abl(n, neumann, pad, 2.5, code)
    np.testing.assert_allclose(g, r, rtol=1e-14, atol=1e-14)


def test_abl_face_mask_ignored_in_2d():
    n, pad = (12, 10), (3, 4)
    g = ha.getABL(n, False, pad, 2.5, np.array([[0, 1], [1, 0]], dtype=bool))
    np.testing.assert_array_equal(g, ha.getABL(n, False, pad, 2.5))
