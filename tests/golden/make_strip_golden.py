"""Writes tests/golden/strip_cpu_64.npz from the reference's own parallel
'strip' projector (CParallelBeamStripKernelProjector2D), on the inputs of
linear_cpu_64.npz (64 x 64 image, 45 angles over [0, pi), 96 detectors of
width 1.0): its forward projection of `vol` and back projection of
`sino_in`.

Run where the reference sources are present and `make -C oracle ref` has
built oracle/_ref/libastra_ref_cpu.so.  The fixture holds outputs only.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import cpu  # noqa: E402


def main():
    strip = cpu.reference_strip()
    if strip is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    fp, bp = strip
    g = np.load(os.path.join(HERE, "linear_cpu_64.npz"))
    out = os.path.join(HERE, "strip_cpu_64.npz")
    np.savez_compressed(out, fp=fp(g["vol"], g["angles"], int(g["n_det"])),
                        bp=bp(g["sino_in"], g["angles"], (64, 64)))
    print("wrote", out)


if __name__ == "__main__":
    main()
