"""Copies the reference's Marshak-wave golden profiles into tests/golden/.

The files are data held by the reference's own regression examples:
examples/marshak_wave_N/reference.txt (512 rows of `x[cm] Tgas[K] Trad[K]`,
radiation-diffusion profiles at t = 1 ns, read by
examples/marshak_wave/MarshakCommon.hpp:175-190 and compared with a relative
L1 below 0.10, :404-413).  Run in a container that has the reference tree;
the committed copies are what the tests read.
"""
from __future__ import annotations

import os
import shutil
import sys

REF = "/root/reference/examples"
HERE = os.path.dirname(os.path.abspath(__file__))


def main() -> int:
    for n in (1, 2, 3, 4):
        src = os.path.join(REF, f"marshak_wave_{n}", "reference.txt")
        dst = os.path.join(HERE, f"marshak_wave_{n}_reference.txt")
        shutil.copyfile(src, dst)
        print(dst)
    return 0


if __name__ == "__main__":
    sys.exit(main())
