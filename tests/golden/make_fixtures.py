"""Copies the reference's golden data for this path into tests/golden/.

The files are data held by the reference's own regression examples:
  * examples/marshak_wave_N/reference.txt (N = 1..4): 512 rows of
    `x[cm] Tgas[K] Trad[K]`, radiation-diffusion profiles at t = 1 ns, read by
    examples/marshak_wave/MarshakCommon.hpp:175-190 and compared with a
    relative L1 below 0.10 (:404-413);
  * examples/densmore2012/data/densmore2012_fig4_mc.csv: 36 points
    `x[cm], T_gas[keV]` digitised from Densmore et al. (2012) Fig. 4, read by
    examples/densmore2012/main.cpp:85-109 and compared with an absolute L1
    below 0.10 keV (:321-354).
Run in a container that has the reference tree; the committed copies are
what the tests read.
"""
from __future__ import annotations

import os
import shutil
import sys

REF = "/root/reference/examples"
HERE = os.path.dirname(os.path.abspath(__file__))


def main() -> int:
    pairs = [(os.path.join(REF, f"marshak_wave_{n}", "reference.txt"),
              os.path.join(HERE, f"marshak_wave_{n}_reference.txt")) for n in (1, 2, 3, 4)]
    pairs.append((os.path.join(REF, "densmore2012", "data", "densmore2012_fig4_mc.csv"),
                  os.path.join(HERE, "densmore2012_fig4_mc.csv")))
    for src, dst in pairs:
        shutil.copyfile(src, dst)
        print(dst)
    return 0


if __name__ == "__main__":
    sys.exit(main())
