"""The GPU 2D FP against the exact line integrals of the linearly
interpolated image (the reference's kernel test,
tests/python/functional/test_line2d.py:399-513): rectangles of ones in
random volume windows, random parallel / fan / vector geometries with
non-unit and oblique detectors.  The reference's own CUDA projector passes
at 2e-2 (8-bit texture weights); the fp32 kernels here meet 5e-4."""
import numpy as np
import pytest

import line2d_known_answer as K

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["parallel", "parallel_vec", "fanflat", "fanflat_vec"])
def test_gpu_fp_matches_linear_kernel_line_integrals(kind):
    import astra
    rng = np.random.default_rng(124)
    for _ in range(25):
        vol, proj, data, rect = K.random_case(kind, rng)
        vg = astra.create_vol_geom(*vol)
        pg = astra.create_proj_geom(*proj)
        pid = astra.create_projector("cuda", pg, vg)
        sid, got = astra.create_sino(data, pid)
        astra.data2d.delete(sid)
        astra.projector.delete(pid)
        shape = astra.functions.geom_size(pg)
        ok, nv, _ = K.oracle_vectors(proj, vg)
        vert = K.fan_kernel_vertical(nv, shape[1]) if ok == "fan" else None
        ref = K.linear_kernel_sinogram(proj, rect, shape, vertical=vert)
        assert np.all(np.isfinite(got))
        assert np.abs(got - ref).max() <= 5e-4
