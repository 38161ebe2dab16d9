"""Host-side logic of the astra mirror (CPU only, no GPU calls)."""
import numpy as np
import pytest

import astra
from astra import geometry as G
from oracle import geometry as OG


def test_create_vol_geom_forms():
    g = astra.create_vol_geom(8)
    assert (g["GridRowCount"], g["GridColCount"]) == (8, 8)
    assert g["option"]["WindowMinX"] == -4.0 and g["option"]["WindowMaxY"] == 4.0
    g = astra.create_vol_geom((6, 10))
    assert (g["GridRowCount"], g["GridColCount"]) == (6, 10)
    g = astra.create_vol_geom(6, 10, 4)
    assert g["GridSliceCount"] == 4 and g["option"]["WindowMinZ"] == -2.0
    g = astra.create_vol_geom(6, 10, -1, 2, -3, 4)
    assert g["option"] == {"WindowMinX": -1, "WindowMaxX": 2, "WindowMinY": -3, "WindowMaxY": 4}
    g = astra.create_vol_geom(2, 3, 4, 0, 1, 0, 2, 0, 3)
    assert g["option"]["WindowMaxZ"] == 3


def test_create_proj_geom_forms():
    a = np.linspace(0, np.pi, 5)
    g = astra.create_proj_geom("parallel", 1.0, 10, a)
    assert g["type"] == "parallel" and g["DetectorCount"] == 10
    g = astra.create_proj_geom("cone", 1.0, 2.0, 8, 9, a, 100, 50)
    assert g["DetectorRowCount"] == 8 and g["DetectorColCount"] == 9
    assert g["DistanceOriginSource"] == 100
    with pytest.raises(astra.AstraError):
        astra.create_proj_geom("cone_vec", 4, 4, np.zeros((3, 11)))
    with pytest.raises(astra.AstraError):
        astra.create_proj_geom("parallel", 1.0)
    with pytest.raises(astra.AstraError):
        astra.create_proj_geom("bogus", 1)


@pytest.mark.parametrize("kind", ["parallel", "fanflat", "parallel_vec", "fanflat_vec"])
def test_geometry_2d_bit_identical_to_oracle(kind):
    rng = np.random.default_rng(0)
    a = rng.random(13) * 2 * np.pi
    vg = astra.create_vol_geom(20, 20, -7, 13, -12, 8)
    if kind == "parallel":
        pg = astra.create_proj_geom(kind, 0.75, 30, a)
        ov = OG.gen_par_projections(a, 30, 0.75)
        okind = "par"
    elif kind == "fanflat":
        pg = astra.create_proj_geom(kind, 1.25, 30, a, 300.0, 120.0)
        ov = OG.gen_fan_projections(a, 30, 300.0, 120.0, 1.25)
        okind = "fan"
    else:
        vecs = rng.normal(size=(13, 6)) * [1, 1, 50, 50, 1, 1]
        pg = astra.create_proj_geom(kind, 30, vecs)
        ov = OG.fanvec_from_astra(vecs, 30)
        okind = "par" if kind == "parallel_vec" else "fan"
    k, v = G.vectors_2d(pg)
    v, osc = G.normalise_2d(k, v, vg)
    ovn, oosc = OG.normalise_2d(ov, okind, vg)
    assert k == okind
    assert np.array_equal(v.astype(np.float64), ovn) and osc == oosc


@pytest.mark.parametrize("kind", ["cone", "parallel3d", "cone_vec", "parallel3d_vec"])
def test_geometry_3d_bit_identical_to_oracle(kind):
    rng = np.random.default_rng(1)
    a = rng.random(11) * 2 * np.pi
    vg = astra.create_vol_geom(10, 12, 14, -6, 6, -4, 6, -7, 7)
    if kind == "cone":
        pg = astra.create_proj_geom(kind, 1.1, 0.9, 8, 9, a, 250.0, 90.0)
        ov = OG.gen_cone_projections(a, 9, 8, 250.0, 90.0, 1.1, 0.9)
    elif kind == "parallel3d":
        pg = astra.create_proj_geom(kind, 1.1, 0.9, 8, 9, a)
        ov = OG.gen_par3d_projections(a, 9, 8, 1.1, 0.9)
    else:
        vecs = rng.normal(size=(11, 12)) * 20
        pg = astra.create_proj_geom(kind, 8, 9, vecs)
        ov = OG.conevec_from_astra(vecs, 9, 8)
    okind = "cone" if kind.startswith("cone") else "par3d"
    k, v = G.vectors_3d(pg)
    v, vs = G.normalise_3d(k, v, vg)
    ovn, ovs = OG.normalise_3d(ov, okind, vg)
    assert k == okind and vs == ovs
    assert np.allclose(v, ovn, rtol=0, atol=1e-12)


def test_geom_2vec_roundtrip():
    a = np.linspace(0, 2 * np.pi, 7, endpoint=False)
    pg = astra.create_proj_geom("cone", 1.0, 1.5, 6, 5, a, 100.0, 40.0)
    vg = astra.create_vol_geom(4, 4, 4)
    vpg = astra.geom_2vec(pg)
    assert vpg["type"] == "cone_vec"
    k1, v1 = G.vectors_3d(pg)
    k2, v2 = G.vectors_3d(vpg)
    assert np.allclose(v1, v2, atol=1e-12)


def test_data_objects_and_dlpack_link():
    vg = astra.create_vol_geom(4, 5)
    i = astra.data2d.create("-vol", vg, 3.0)
    assert np.all(astra.data2d.get(i) == 3.0) and astra.data2d.get(i).shape == (4, 5)
    astra.data2d.store(i, np.arange(20, dtype=np.float32).reshape(4, 5))
    assert astra.data2d.get(i)[3, 4] == 19
    with pytest.raises(astra.AstraError):
        astra.data2d.store(i, np.zeros((5, 4)))
    # link shares memory (zero copy)
    arr = np.zeros((4, 5), dtype=np.float32)
    j = astra.data2d.link("-vol", vg, arr)
    astra.data2d.store(j, 2.0)
    assert np.all(arr == 2.0)
    # DLPack contract: read-only and non-contiguous are rejected (ValueError),
    # objects without __dlpack__ are a TypeError, wrong dtype a ValueError
    ro = np.zeros((4, 5), dtype=np.float32)
    ro.setflags(write=False)
    with pytest.raises(ValueError):
        astra.data2d.link("-vol", vg, ro)
    with pytest.raises(ValueError):
        astra.data2d.link("-vol", vg, np.zeros((5, 4), dtype=np.float32).T)
    with pytest.raises(ValueError):
        astra.data2d.link("-vol", vg, np.zeros((4, 5), dtype=np.float64))
    with pytest.raises(TypeError):
        astra.data2d.link("-vol", vg, [[0.0] * 5] * 4)
    astra.data2d.delete([i, j])
    with pytest.raises(astra.AstraError):
        astra.data2d.get(i)


def test_data3d_shapes():
    vg = astra.create_vol_geom(4, 5, 6)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, 7, 8, np.zeros(3), 10.0, 10.0)
    v = astra.data3d.create("-vol", vg)
    p = astra.data3d.create("-sino", pg)
    assert astra.data3d.dimensions(v) == (6, 4, 5)
    assert astra.data3d.dimensions(p) == (7, 3, 8)
    astra.data3d.delete([v, p])


def test_algorithm_config_errors():
    vg = astra.create_vol_geom(4, 4)
    pg = astra.create_proj_geom("parallel", 1.0, 6, np.zeros(3))
    s = astra.data2d.create("-sino", pg)
    r = astra.data2d.create("-vol", vg)
    with pytest.raises(astra.AstraError):
        astra.algorithm.create({"type": "NOT_AN_ALGORITHM"})
    with pytest.raises(astra.AstraError):
        astra.algorithm.create({"type": "ART", "ProjectionDataId": s, "ReconstructionDataId": r})
    with pytest.raises(astra.AstraError):
        astra.algorithm.create({"type": "SART_CUDA", "ProjectionDataId": s, "ReconstructionDataId": r,
                                "option": {"ProjectionOrder": "backwards"}})
    m = astra.data2d.create("-vol", vg, 1.0)
    with pytest.raises(astra.AstraError):
        astra.algorithm.create({"type": "EM_CUDA", "ProjectionDataId": s, "ReconstructionDataId": r,
                                "option": {"ReconstructionMaskId": m}})
    sart = astra.algorithm.create({"type": "SART_CUDA", "ProjectionDataId": s, "ReconstructionDataId": r,
                                   "option": {"ProjectionOrder": "custom", "ProjectionOrderList": [2, 0, 1]}})
    assert astra.algorithm._reg.get(sart).order == [2, 0, 1]
    astra.algorithm.delete(sart)
    with pytest.raises(astra.AstraError):
        astra.create_projector("linear", pg, vg)
    cfg = astra.astra_dict("SIRT_CUDA")
    cfg.update(ProjectionDataId=s, ReconstructionDataId=r)
    cfg["option"] = {"MinConstraint": 0.0, "UseMaxConstraint": "yes", "MaxConstraintValue": 2.0}
    a = astra.algorithm.create(cfg)
    obj = astra.algorithm._reg.get(a)
    assert obj.options == {"Relaxation": 1.0, "MinConstraint": 0.0, "MaxConstraint": 2.0}
    astra.algorithm.delete(a)
    astra.data2d.delete([s, r])


def test_add_noise_to_sino_statistics():
    sino = np.full((50, 80), 2.0, dtype=np.float32)
    noisy = astra.add_noise_to_sino(sino, 1e5, seed=3)
    # b = -max log(Poisson(I0 e^{-1}) / I0): mean ~2, sd ~ 2 / sqrt(I0 / e)
    assert abs(noisy.mean() - 2.0) < 1e-3
    assert 0.5 < noisy.std() / (2.0 / np.sqrt(1e5 / np.e)) < 1.5


def test_cylcone_geometry_matches_oracle_translation():
    rng = np.random.default_rng(5)
    vecs = np.c_[rng.normal(size=(7, 12)) * 20, np.full(7, 80.0)]
    vg = astra.create_vol_geom(10, 12, 14, -6, 6, -4, 6, -7, 7)
    pg = astra.create_proj_geom("cyl_cone_vec", 9, 8, vecs)
    k, v = G.vectors_3d(pg)
    assert k == "cylcone"
    v, vs = G.normalise_3d(k, v, vg)
    ov, ovs = OG.translate_cylcone(vecs, vg)
    assert np.array_equal(v, ov) and vs == ovs
    with pytest.raises(astra.AstraError):
        G.vectors_3d(astra.create_proj_geom("cyl_cone_vec", 9, 8, np.c_[vecs[:, :12], -np.ones(7)]))
    with pytest.raises(astra.AstraError):
        astra.create_proj_geom("cyl_cone_vec", 9, 8, vecs[:, :12])


def test_add_noise_to_sino_matches_reference_draw_order():
    # the reference draws one np.random.poisson per element in nested loops
    # from the seeded global legacy generator and restores its state
    # (python/astra/functions.py:94-119)
    rng = np.random.default_rng(9)
    sino = (rng.random((13, 17)) * 3).astype(np.float32)
    np.random.seed(1234)
    before = np.random.random()
    np.random.seed(1234)
    got = astra.add_noise_to_sino(sino, 1e4, seed=77)
    assert np.random.random() == before  # global state restored
    np.random.seed(77)
    ct = 1e4 * np.exp(-(sino / sino.max()))
    c = np.zeros_like(ct)
    for i in range(c.shape[0]):
        for j in range(c.shape[1]):
            c[i, j] = np.random.poisson(ct[i, j])
    ref = -sino.max() * np.log(c / 1e4)
    assert got.dtype == ref.dtype and np.array_equal(got, ref)
