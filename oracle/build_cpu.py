"""Build recipe for the oracle's C restatement (oracle/slmg_cpu.c).

Test infrastructure only: the library is the CPU baseline and a second check
of the NumPy oracle; the product never loads it.  Output: oracle/libslmg_cpu.so.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "slmg_cpu.c")
OUT = os.path.join(HERE, "libslmg_cpu.so")


def build() -> str:
    cmd = ["gcc", "-O3", "-std=gnu11", "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wno-unused-result",
           SRC, "-o", OUT + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build())
