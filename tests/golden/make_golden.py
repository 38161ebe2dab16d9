"""Writes tests/golden/linear_cpu_64.npz from the reference's own CPU library.

Run in a checkout where the reference sources are present and
`make -C oracle ref` has built oracle/_ref/libastra_ref_cpu.so.  The fixture
holds inputs and outputs only (data, not code): the reference's
CParallelBeamLinearKernelProjector2D forward and back projection and 10
iterations of CSirtAlgorithm at 64 x 64, 45 angles over [0, pi), 96
detectors of width 1.0.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "astra-toolbox_amd"))

from astra import phantom  # noqa: E402
from oracle import cpu  # noqa: E402


def main():
    ref = cpu.reference()
    if ref is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    n, na, nd = 64, 45, 96
    angles = np.linspace(0, np.pi, na, endpoint=False).astype(np.float32)
    vol = phantom.ellipses_2d(n, seed=1)
    rng = np.random.default_rng(0)
    sino_in = rng.random((na, nd), dtype=np.float32)
    fp = ref.fp(vol, angles, nd)
    bp = ref.bp(sino_in, angles, (n, n))
    sirt10 = ref.sirt(fp, angles, (n, n), 10)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "linear_cpu_64.npz")
    np.savez_compressed(out, vol=vol, angles=angles, sino_in=sino_in, fp=fp, bp=bp, sirt10=sirt10,
                        n_det=np.int32(nd), det_width=np.float32(1.0))
    print("wrote", out)


if __name__ == "__main__":
    main()
