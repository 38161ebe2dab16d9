"""The job splitter of astra.composite (CCompositeGeometryManager::splitJobs,
src/CompositeGeometryManager.cpp:218-379) on the host: the parts tile the
data, each part fits the budget, and the sub-geometries tile the volume
window and the projection set."""
import numpy as np
import pytest

import astra
from astra import composite as C
from astra import geometry as G


@pytest.mark.parametrize("seed", range(6))
def test_fp_bp_plans_tile_and_fit(seed):
    rng = np.random.default_rng(seed)
    vshape = tuple(int(x) for x in rng.integers(8, 300, 3))
    pshape = tuple(int(x) for x in rng.integers(8, 400, 3))
    full = C._fp_scratch(vshape) + C._nbytes(vshape) + C._nbytes(pshape)
    cap = int(full * rng.uniform(0.2, 0.9))
    plan = C.plan_fp(vshape, pshape, cap)
    nv, na, nu = pshape
    nz, ny, nx = vshape
    assert plan is not None
    angles = [r for r, _ in plan]
    assert angles[0][0] == 0 and angles[-1][1] == na
    assert all(a[1] == b[0] for a, b in zip(angles, angles[1:]))
    for (a0, a1), slabs in plan:
        assert slabs[0][0] == 0 and slabs[-1][1] == nz
        assert all(a[1] == b[0] for a, b in zip(slabs, slabs[1:]))
        for z0, z1 in slabs:
            use = (a1 - a0) * nv * nu * 4 + C._nbytes((z1 - z0, ny, nx)) + C._fp_scratch((z1 - z0, ny, nx))
            assert use <= cap
    full = C._bp_scratch(pshape) + C._nbytes(vshape) + C._nbytes(pshape)
    cap = int(full * rng.uniform(0.2, 0.9))
    plan = C.plan_bp(vshape, pshape, cap)
    for (z0, z1), chunks in plan:
        assert chunks[0][0] == 0 and chunks[-1][1] == na
        for a0, a1 in chunks:
            use = C._nbytes((z1 - z0, ny, nx)) + C._nbytes((nv, a1 - a0, nu)) + \
                C._bp_scratch((nv, a1 - a0, nu))
            assert use <= cap
    assert C.plan_fp(vshape, pshape, 10 ** 13) is None and C.plan_bp(vshape, pshape, 10 ** 13) is None


def test_sub_geometries():
    vg = astra.create_vol_geom(8, 9, 10, -5, 4, -3, 5, -2.5, 7.5)
    parts = [C.vol_slab_geom(vg, z0, z1) for z0, z1 in ((0, 3), (3, 7), (7, 10))]
    assert [p["GridSliceCount"] for p in parts] == [3, 4, 3]
    assert parts[0]["option"]["WindowMinZ"] == -2.5 and parts[-1]["option"]["WindowMaxZ"] == 7.5
    for a, b in zip(parts, parts[1:]):
        assert a["option"]["WindowMaxZ"] == b["option"]["WindowMinZ"]
    for p in parts:  # voxel size unchanged
        assert G.pixel_lengths(p) == G.pixel_lengths(vg)
    pg = astra.create_proj_geom("cone", 1.0, 1.2, 6, 7, np.linspace(0, 3, 9), 100.0, 40.0)
    _, full = G.vectors_3d(pg)
    sub = C.proj_angle_geom(pg, 2, 6)
    assert sub["type"] == "cone_vec" and sub["Vectors"].shape == (4, 12)
    _, part = G.vectors_3d(sub)
    assert np.allclose(part, full[2:6], rtol=0, atol=1e-12)


def test_set_gpu_index_memory_floor():
    with pytest.raises(astra.AstraError):
        astra.set_gpu_index(0, memory=1000)
    astra.set_gpu_index(0, memory=0)


def test_vertical_reduction_closed_forms():
    """parallel3d with unit pixels and voxels: detector row v sees height
    z = v - (rows - 1) / 2, so the rows a slab reads and the slices that
    reach the detector follow from the window, the half-voxel growth and
    the one-row / one-slice margins (src/CompositeGeometryManager.cpp:
    382-410, 526-625, 854-866)."""
    pg = astra.create_proj_geom("parallel3d", 1.0, 1.0, 40, 64, np.linspace(0, np.pi, 30, endpoint=False))
    vg = astra.create_vol_geom(32, 32, 100)                    # z window [-50, 50]
    # slab [50, 60): z in [0, 10], grown to [-0.5, 10.5] -> v in [19, 30]
    assert C.rows_needed(C.vol_slab_geom(vg, 50, 60), pg) == (18, 31)
    assert C.rows_needed(C.vol_slab_geom(vg, 0, 10), pg) == (0, 0)     # z <= -40: off the detector
    assert C.rows_needed(vg, pg) == (0, 40)
    # rows [0, 40) see z in [-19.5, 19.5] (+-1 row: [-20.5, 20.5]); slice k
    # (centre k - 49.5) reaches it from k = 29 to 70, plus one slice each side
    z0, z1 = C.slices_needed(vg, pg)
    assert (z0, z1) == (28, 72)
    assert C.slices_needed(astra.create_vol_geom(32, 32, 20), pg) == (0, 20)


def test_row_subgeometry_keeps_pixels():
    """Rows [r0, r1) of a detector are the same pixels: a point projects to
    row v in the full geometry and to v - r0 in the sub-geometry."""
    from astra import geometry as G
    pg = astra.create_proj_geom("cone", 1.0, 1.3, 30, 40, np.linspace(0, 2 * np.pi, 7), 150.0, 60.0)
    sub = C.proj_rows_geom(pg, 6, 19)
    assert sub["DetectorRowCount"] == 13
    pts = np.random.default_rng(0).uniform(-10, 10, (20, 3))
    u, v = G.project_points_3d(pg, pts)
    su, sv = G.project_points_3d(sub, pts)
    assert np.allclose(su, u, atol=1e-5) and np.allclose(sv, v - 6, atol=1e-5)
