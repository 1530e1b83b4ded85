"""Builds libmycel_hip.so for gfx950 with hipcc (python -m mycel_amd.build)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmycel_hip.so")


def build(force: bool = True, jobs: int = 6) -> str:
    cmd = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    if not os.path.exists(LIB):
        raise RuntimeError("hipcc did not produce " + LIB)
    return LIB


if __name__ == "__main__":
    build(force="--no-force" not in sys.argv)
    print(LIB)
