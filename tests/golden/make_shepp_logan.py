"""Writes tests/golden/shepp_logan.npz from the reference's own
generateSheppLogan / generateSheppLogan3D (src/SheppLogan.cpp:187-246).

Run where the reference sources are present and `make -C oracle ref` has
built oracle/_ref/libastra_ref_cpu.so.  The fixture holds the generated
phantoms only (data): 2D 50 x 60 and 3 x 2, 3D 50 x 40 x 30 and 5 x 4 x 3,
original and modified densities.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import cpu  # noqa: E402

SHAPES = {"2d_50x60": (50, 60), "2d_3x2": (3, 2), "3d_50x40x30": (50, 40, 30), "3d_5x4x3": (5, 4, 3)}


def main():
    out = {}
    for name, shape in SHAPES.items():
        for modified in (True, False):
            a = cpu.reference_shepp_logan(shape, modified)
            if a is None:
                raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
            out["%s_%s" % (name, "mod" if modified else "orig")] = a
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shepp_logan.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
