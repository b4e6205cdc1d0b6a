"""Writes tests/golden/seg_salt_2d.npz from the reference's data file
examples/SEGmodel2Dsalt.dat (a 2-D slice of the SEG/EAGE salt model: 128
depth rows x 256 columns of P velocity in m/s, whole numbers 1500..4480),
the medium that the commented-out line of test/ShiftedLaplacianTest.jl:14
reads.  Stored as uint16 in the file's own row order (rows = depth), with
the SHA-256 of the float64 array for the loader's check.

Run (in a checkout next to the reference): python tests/golden/make_seg_fixture.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/examples/SEGmodel2Dsalt.dat"


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    v = np.loadtxt(src)
    assert v.shape == (128, 256) and np.all(v == np.round(v)) and v.min() >= 0 and v.max() < 65536
    sha = hashlib.sha256(np.ascontiguousarray(v, dtype=np.float64).tobytes()).hexdigest()
    np.savez_compressed(os.path.join(HERE, "seg_salt_2d.npz"), v=v.astype(np.uint16), sha256=np.array(sha))
    print(v.shape, v.min(), v.max(), sha)


if __name__ == "__main__":
    main()
