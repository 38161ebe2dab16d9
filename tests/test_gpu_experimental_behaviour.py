"""astra.experimental / projector(3d).direct_* on the reference's unit-test
shapes (tests/python/unit/test_experimental.py:86-155,
tests/python/unit/test_experimental_2d.py:81-112): a shifted 30 x 40 x 50
volume, a 20 x 45 detector, 180 angles; a 60 x 50 image, 40 bins.

The reference checks that outputs are non-zero and that MODE_ADD changes
them; here the same calls are held to what the arithmetic fixes exactly:
ADD after SET doubles the output bit for bit (the projectors are
deterministic), a composite job over two volumes / projection sets is twice
the single job, accumulate_FDK twice is twice once, and the direct calls
equal the data-object paths (create_sino / create_backprojection).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ANG = np.linspace(0, 2 * np.pi, 180, endpoint=False)
MODE_ADD, MODE_SET = 0, 1


def _astra():
    import astra
    import astra.experimental
    return astra


def _vg3():
    return _astra().create_vol_geom(40, 30, 50, -14.0, 16.0, -18.0, 22.0, -22.0, 28.0)


def _pg3(kind):
    a = _astra()
    if kind.startswith("parallel3d"):
        g = a.create_proj_geom("parallel3d", 1.0, 1.0, 20, 45, ANG)
    else:
        g = a.create_proj_geom("cone", 1.0, 1.0, 20, 45, ANG, 100.0, 100.0)
    if kind == "cyl_cone_vec":
        v = a.geom_2vec(g)["Vectors"]
        return a.create_proj_geom("cyl_cone_vec", 20, 45, np.hstack([v, np.full((180, 1), 200.0)]))
    return a.geom_2vec(g) if kind.endswith("_vec") else g


KINDS3 = ["parallel3d", "parallel3d_vec", "cone", "cone_vec", "cyl_cone_vec"]


@pytest.fixture
def proj3(request):
    a = _astra()
    pg = _pg3(request.param)
    pid = a.create_projector("cuda3d", pg, _vg3())
    yield a, pg, pid
    a.projector3d.delete(pid)


@pytest.mark.parametrize("proj3", KINDS3, indirect=True)
def test_direct_fpbp3d_set_then_add(proj3):
    a, pg, pid = proj3
    vol = np.ones((50, 40, 30), np.float32)
    proj = np.zeros((20, 180, 45), np.float32)
    a.experimental.direct_FPBP3D(pid, vol, proj, MODE_SET, "FP")
    twice = proj.copy()
    a.experimental.direct_FPBP3D(pid, vol, twice, MODE_ADD, "FP")
    assert proj.max() > 0 and np.array_equal(twice, 2 * proj)
    back = np.zeros((50, 40, 30), np.float32)
    ones = np.ones((20, 180, 45), np.float32)
    a.experimental.direct_FPBP3D(pid, back, ones, MODE_SET, "BP")
    back2 = back.copy()
    a.experimental.direct_FPBP3D(pid, back2, ones, MODE_ADD, "BP")
    assert back.max() > 0 and np.array_equal(back2, 2 * back)
    # projector3d.direct_FP / direct_BP (out=) are the same jobs
    out = np.full((20, 180, 45), 9.0, np.float32)
    assert a.projector3d.direct_FP(pid, vol, out=out) is out and np.array_equal(out, proj)
    outb = np.full((50, 40, 30), 9.0, np.float32)
    a.projector3d.direct_BP(pid, ones, out=outb)
    assert np.array_equal(outb, back)


@pytest.mark.parametrize("proj3", KINDS3, indirect=True)
def test_composite_and_accumulate(proj3):
    a, pg, pid = proj3
    vg = _vg3()
    vol = np.ones((50, 40, 30), np.float32)
    single = np.zeros((20, 180, 45), np.float32)
    a.experimental.direct_FP3D(pid, vol, single)
    vids = [a.data3d.create("-vol", vg, vol) for _ in range(2)]
    pids = [a.data3d.create("-sino", pg, 0.0) for _ in range(2)]
    a.experimental.do_composite(pid, vids, pids, MODE_SET, "FP")
    first = [a.data3d.get(i) for i in pids]
    a.experimental.do_composite(pid, vids, pids, MODE_ADD, "FP")
    second = [a.data3d.get(i) for i in pids]
    for f, s in zip(first, second):
        assert np.array_equal(f, 2 * single) and np.array_equal(s, 4 * single)
    a.experimental.accumulate_FP(pid, vids[0], pids[0])
    assert np.array_equal(a.data3d.get(pids[0]), 5 * single)
    # BP side: each volume receives both projection sets
    ones = np.ones((20, 180, 45), np.float32)
    bsingle = np.zeros((50, 40, 30), np.float32)
    a.experimental.direct_BP3D(pid, bsingle, ones)
    for i in pids:
        a.data3d.store(i, ones)
    a.experimental.do_composite_BP(pid, vids, pids)
    for i in vids:
        assert np.array_equal(a.data3d.get(i), 2 * bsingle)
    a.experimental.accumulate_BP(pid, vids[1], pids[1])
    assert np.array_equal(a.data3d.get(vids[1]), 3 * bsingle)
    a.data3d.delete(vids + pids)


@pytest.mark.parametrize("kind", ["cone", "cone_vec"])
def test_accumulate_fdk_linked_host_buffers(kind):
    a = _astra()
    pg = _pg3(kind)
    pid = a.create_projector("cuda3d", pg, _vg3())
    proj = np.ones((20, 180, 45), np.float32)
    vol = np.zeros((50, 40, 30), np.float32)
    sid = a.data3d.link("-sino", pg, proj)
    vid = a.data3d.link("-vol", _vg3(), vol)
    a.experimental.accumulate_FDK(pid, vid, sid)
    once = vol.copy()
    a.experimental.accumulate_FDK(pid, vid, sid)
    assert np.abs(once).max() > 0
    assert np.array_equal(vol, 2 * once)
    assert np.array_equal(proj, np.ones_like(proj))  # input left unchanged
    a.data3d.delete([sid, vid])
    a.projector3d.delete(pid)


@pytest.mark.parametrize("kind", KINDS3)
def test_geometry_helpers_take_keywords(kind):
    a = _astra()
    pg = _pg3(kind)
    vmin, vmax = a.experimental.getProjectedBBox(pg, minx=0, maxx=1, miny=2, maxy=3, minz=4, maxz=5)
    u, v = a.experimental.projectPoint(pg, x=0, y=1, z=2, angle=3)
    assert 0 <= u <= 44 and 0 <= v <= 19
    _, vin = a.experimental.projectPoint(pg, x=0.5, y=2.5, z=4.5, angle=7)
    assert vmin <= vin <= vmax


# --------------------------------------------------------------- 2D ------
def _vg2():
    return _astra().create_vol_geom(50, 60, -29.0, 31.0, -23.0, 27.0)


@pytest.mark.parametrize("kind", ["parallel", "parallel_vec", "fanflat", "fanflat_vec"])
def test_direct_2d_equal_data_object_paths(kind):
    a = _astra()
    if kind.startswith("parallel"):
        pg = a.create_proj_geom("parallel", 1.0, 40, ANG)
    else:
        pg = a.create_proj_geom("fanflat", 1.0, 40, ANG, 100.0, 100.0)
    if kind.endswith("_vec"):
        pg = a.geom_2vec(pg)
    pid = a.create_projector("cuda", pg, _vg2())
    vol = np.ones((50, 60), np.float32)
    sino = np.zeros((180, 40), np.float32)
    a.experimental.direct_FP2D(pid, vol, sino)
    sid, ref = a.create_sino(vol, pid)
    a.data2d.delete(sid)
    assert sino.max() > 0 and np.array_equal(sino, ref)
    out = np.zeros((180, 40), np.float32)
    a.projector.direct_FP(pid, vol, out=out)
    assert np.array_equal(out, ref)
    ones = np.ones((180, 40), np.float32)
    back = np.zeros((50, 60), np.float32)
    a.experimental.direct_BP2D(pid, back, ones)
    bid, bref = a.create_backprojection(ones, pid)
    a.data2d.delete(bid)
    assert back.max() > 0 and np.array_equal(back, bref)
    outb = np.zeros((50, 60), np.float32)
    a.projector.direct_BP(pid, ones, out=outb)
    assert np.array_equal(outb, bref)
    a.projector.delete(pid)
