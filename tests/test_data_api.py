"""Data-object behaviour of astra.data2d / astra.data3d as the reference's
unit tests state it (tests/python/unit/test_data_2d.py,
tests/python/unit/test_data_3d.py), on the host: initialisers and dtypes,
linking and shared views, dimensions, geometry round trips (a volume's
window comes back under 'options'), change_geometry, delete / clear,
info() listing, and the Shepp-Logan phantoms (bit-equal to the reference's
generator: its golden fixture, and its build when present)."""
import os

import numpy as np
import pytest

import astra

ANG = np.linspace(0, 2 * np.pi, 180, endpoint=False)
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shepp_logan.npz")


def _geom(mod, kind):
    if mod is astra.data3d:
        if kind == "-vol":
            return astra.create_vol_geom(40, 30, 50)
        return astra.create_proj_geom("parallel3d", 1.0, 1.0, 20, 45, ANG)
    if kind == "-vol":
        return astra.create_vol_geom(40, 30)
    return astra.create_proj_geom("parallel", 1.0, 45, ANG)


def _shape(mod, kind):
    if mod is astra.data3d:
        return (50, 40, 30) if kind == "-vol" else (20, 180, 45)
    return (40, 30) if kind == "-vol" else (180, 45)


MODS = [astra.data2d, astra.data3d]
KINDS = ["-vol", "-sino"]


@pytest.mark.parametrize("mod", MODS, ids=["2d", "3d"])
@pytest.mark.parametrize("kind", KINDS)
def test_initialisers_and_dtypes(mod, kind):
    g = _geom(mod, kind)
    i = mod.create(kind, g)
    assert mod.get(i).shape == _shape(mod, kind) and not mod.get(i).any()
    j = mod.create(kind, g, 1.0)
    assert np.all(mod.get(j) == 1.0)
    rng = np.random.default_rng(0)
    for dt in (np.float32, np.float64, np.int32, np.uint8, bool):
        m = (rng.random(_shape(mod, kind)) * 10).astype(dt)
        k = mod.create(kind, g, m)
        got = mod.get(k)
        assert got.dtype == np.float32 and np.array_equal(got, m.astype(np.float32))
        mod.delete(k)
    assert mod.dimensions(i) == _shape(mod, kind)
    mod.delete([i, j])


@pytest.mark.parametrize("mod", MODS, ids=["2d", "3d"])
@pytest.mark.parametrize("kind", KINDS)
def test_link_and_shared_views(mod, kind):
    g = _geom(mod, kind)
    arr = np.random.default_rng(1).random(_shape(mod, kind)).astype(np.float32)
    i = mod.link(kind, g, arr)
    assert np.array_equal(mod.get(i), arr)
    mod.store(i, 2.0)
    assert np.all(arr == 2.0)                      # the linked array is the storage
    j = mod.create(kind, g, 0.0)
    shared = mod.get_shared(j)
    shared[...] = 3.0
    assert np.all(mod.get(j) == 3.0)
    with pytest.raises(Exception):
        mod.link(kind, g, arr.astype(np.float64))  # float32 only
    mod.delete([i, j])


@pytest.mark.parametrize("mod", MODS, ids=["2d", "3d"])
def test_geometry_round_trip_and_change(mod):
    gv = _geom(mod, "-vol")
    i = mod.create("-vol", gv)
    out = mod.get_geometry(i)
    assert "option" not in out and out["options"] == gv["option"]
    assert {k: v for k, v in out.items() if k != "options"} == \
        {k: v for k, v in gv.items() if k != "option"}
    if mod is astra.data3d:
        new = astra.create_vol_geom(40, 30, 50, 0, 1, 2, 3, 4, 5)
        want = {"WindowMinX": 0, "WindowMaxX": 1, "WindowMinY": 2, "WindowMaxY": 3,
                "WindowMinZ": 4, "WindowMaxZ": 5}
    else:
        new = astra.create_vol_geom(40, 30, 0, 1, 2, 3)
        want = {"WindowMinX": 0, "WindowMaxX": 1, "WindowMinY": 2, "WindowMaxY": 3}
    mod.change_geometry(i, new)
    assert mod.get_geometry(i)["options"] == want
    # a geometry read back ('options') is accepted where one is expected
    j = mod.create("-vol", mod.get_geometry(i))
    assert mod.get_geometry(j) == mod.get_geometry(i)
    gs = _geom(mod, "-sino")
    k = mod.create("-sino", gs)
    back = dict(mod.get_geometry(k))
    assert np.allclose(back.pop("ProjectionAngles"), gs["ProjectionAngles"])
    assert back == {kk: v for kk, v in gs.items() if kk != "ProjectionAngles"}
    if mod is astra.data3d:
        na = np.random.default_rng(2).random(180)
        mod.change_geometry(k, astra.create_proj_geom("parallel3d", 2.0, 3.0, 20, 45, na))
        cg = mod.get_geometry(k)
        assert cg["DetectorSpacingX"] == 2.0 and cg["DetectorSpacingY"] == 3.0
        assert np.allclose(cg["ProjectionAngles"], na)
    mod.delete([i, j, k])


@pytest.mark.parametrize("mod", MODS, ids=["2d", "3d"])
def test_delete_clear_and_info(mod, capsys):
    g = _geom(mod, "-vol")
    mod.clear()
    capsys.readouterr()
    i = mod.create("-vol", g)
    mod.info()
    out = capsys.readouterr().out
    assert len(out.split("\n")) - 5 == 1
    line = out.split("\n")[2]
    dims = "x".join(str(n) for n in reversed(_shape(mod, "-vol")))
    assert line.startswith("%d \tv     %s volume data" % (i, dims))
    mod.delete(i)
    mod.info()
    assert len(capsys.readouterr().out.split("\n")) - 5 == 0
    with pytest.raises(astra.log.AstraError):
        mod.get(i)
    a, b = mod.create("-vol", g), mod.create("-vol", g)
    mod.clear()
    for k in (a, b):
        with pytest.raises(astra.log.AstraError):
            mod.get(k)


@pytest.mark.parametrize("modified", [True, False])
def test_shepp_logan_matches_reference_generator(modified):
    from astra import phantom
    from oracle import cpu
    gold = np.load(GOLDEN)
    tag = "mod" if modified else "orig"
    for name, shape in (("2d_50x60", (50, 60)), ("2d_3x2", (3, 2)), ("3d_50x40x30", (50, 40, 30)),
                        ("3d_5x4x3", (5, 4, 3))):
        fn = phantom.shepp_logan_2d if len(shape) == 2 else phantom.shepp_logan_3d
        assert np.array_equal(fn(*shape, modified=modified), gold["%s_%s" % (name, tag)])
    if cpu.reference_shepp_logan((4, 4)) is None:
        return
    for shape in ((64, 64), (127, 33), (7, 9), (1, 1), (17, 33, 21), (64, 48, 40), (2, 3, 4)):
        fn = phantom.shepp_logan_2d if len(shape) == 2 else phantom.shepp_logan_3d
        assert np.array_equal(fn(*shape, modified=modified), cpu.reference_shepp_logan(shape, modified))


@pytest.mark.parametrize("mod", MODS, ids=["2d", "3d"])
def test_shepp_logan_data_objects(mod):
    g = _geom(mod, "-vol")
    i, d = mod.shepp_logan(g)
    assert d.shape == _shape(mod, "-vol") and d.max() == np.float32(1.0) and np.array_equal(mod.get(i), d)
    j = mod.shepp_logan(g, modified=False, returnData=False)
    assert isinstance(j, int) and mod.get(j).max() == np.float32(2.0)
    mod.delete([i, j])
