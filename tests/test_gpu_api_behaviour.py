"""Behaviour of the algorithm API on the reference's own unit-test shapes.

The reference checks its GPU algorithms through the Python API with
tests/python/unit/test_algorithms_2d.py:141-480 and
tests/python/unit/test_algorithms_3d.py:161-325: every algorithm on every
geometry changes its output, options (supersampling, short scan, filters)
change the result, MinConstraint / MaxConstraint are met exactly, masked
voxels keep their value, SART takes every projection order, and
get_res_norm is positive, repeatable and leaves the reconstruction alone.
The same checks run here on the same shapes (a 60 x 50 image, 40 bins,
180 angles; a shifted 30 x 40 x 50 volume, 20 x 45 detector), written
afresh, plus a few that pin values the reference only checks loosely: the
residual norm equals ||p - FP(x)|| of the state it reports on, and two
calls that start from the same state agree bit for bit.

Algorithms are built without a ProjectorId, as the reference's GPU tests
build them, so the projector comes from the data objects' geometries.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INIT = 1.0
A2 = np.linspace(0, 2 * np.pi, 180, endpoint=False)
A3 = A2


def _astra():
    import astra
    return astra


# --------------------------------------------------------------- 2D ------
def _vg2():
    a = _astra()
    return a.create_vol_geom(50, 60, -30.0, 30.0, -25.0, 25.0)


def _pg2(kind):
    a = _astra()
    if kind in ("parallel", "parallel_vec"):
        g = a.create_proj_geom("parallel", 1.0, 40, A2)
    elif kind in ("fanflat", "fanflat_vec"):
        g = a.create_proj_geom("fanflat", 1.0, 40, A2, 100.0, 100.0)
    elif kind == "short_scan":
        half = np.arctan2(20.0, 200.0)
        g = a.create_proj_geom("fanflat", 1.0, 40, np.linspace(0, np.pi + 2 * half, 180),
                               100.0, 100.0)
    else:
        raise ValueError(kind)
    return a.geom_2vec(g) if kind.endswith("_vec") else g


def _sino2_input():
    """-10 everywhere but the central half of the bins (10): large and
    negative enough to push a few iterations past any box constraint."""
    s = np.full((180, 40), -10.0)
    s[:, 10:30] = 10.0
    return s


def _config2(atype, pg, opts=None):
    a = _astra()
    cfg = a.astra_dict(atype)
    vid = a.data2d.create("-vol", _vg2(), INIT)
    if atype.startswith("FP"):
        cfg["VolumeDataId"] = vid
        cfg["ProjectionDataId"] = a.data2d.create("-sino", pg, INIT)
    else:
        cfg["ReconstructionDataId"] = vid
        cfg["ProjectionDataId"] = a.data2d.create("-sino", pg, _sino2_input())
    if opts:
        cfg["option"] = dict(opts)
    return cfg


def _default_iters(atype):
    if atype.startswith(("SIRT", "CGLS", "EM")):
        return 2
    if atype.startswith("SART"):
        return 180
    return 1


def _output(cfg, iters=None, keep=False):
    """Run a fresh algorithm on cfg's data; return the written output."""
    a = _astra()
    d = a.data3d if "3D" in cfg["type"] or cfg["type"] == "FDK_CUDA" else a.data2d
    aid = a.algorithm.create(cfg)
    a.algorithm.run(aid, _default_iters(cfg["type"]) if iters is None else iters)
    out_id = cfg["ProjectionDataId"] if cfg["type"].startswith("FP") else cfg["ReconstructionDataId"]
    out = d.get(out_id)
    a.algorithm.delete(aid)
    if not keep:
        d.delete([cfg["ProjectionDataId"], out_id])
    return out


@pytest.mark.parametrize("kind", ["parallel", "parallel_vec", "fanflat", "fanflat_vec"])
@pytest.mark.parametrize("atype", ["FP_CUDA", "BP_CUDA", "FBP_CUDA", "SIRT_CUDA", "SART_CUDA",
                                   "CGLS_CUDA", "EM_CUDA"])
def test_2d_algorithm_changes_output(kind, atype):
    out = _output(_config2(atype, _pg2(kind)))
    assert out.shape == ((180, 40) if atype == "FP_CUDA" else (50, 60))
    assert np.all(np.isfinite(out))
    assert not np.allclose(out, INIT)


@pytest.mark.parametrize("kind", ["parallel", "fanflat"])
def test_2d_detector_supersampling_fp(kind):
    pg = _pg2(kind)
    plain = _output(_config2("FP_CUDA", pg))
    ss = _output(_config2("FP_CUDA", pg, {"DetectorSuperSampling": 3}))
    assert not np.allclose(ss, INIT)
    assert not np.allclose(ss, plain)


@pytest.mark.parametrize("kind", ["parallel", "fanflat"])
@pytest.mark.parametrize("atype", ["SIRT_CUDA", "SART_CUDA", "CGLS_CUDA", "EM_CUDA"])
def test_2d_detector_supersampling_iterative(kind, atype):
    pg = _pg2(kind)
    plain = _output(_config2(atype, pg, {"ProjectionOrder": "sequential"} if atype == "SART_CUDA" else None))
    opts = {"DetectorSuperSampling": 3}
    if atype == "SART_CUDA":
        opts["ProjectionOrder"] = "sequential"
    ss = _output(_config2(atype, pg, opts))
    assert not np.allclose(ss, INIT)
    assert not np.allclose(ss, plain)


@pytest.mark.parametrize("kind", ["parallel", "fanflat"])
@pytest.mark.parametrize("atype", ["BP_CUDA", "SIRT_CUDA", "CGLS_CUDA", "EM_CUDA"])
def test_2d_pixel_supersampling(kind, atype):
    pg = _pg2(kind)
    plain = _output(_config2(atype, pg))
    ss = _output(_config2(atype, pg, {"PixelSuperSampling": 3}))
    assert not np.allclose(ss, INIT)
    assert not np.allclose(ss, plain)


@pytest.mark.parametrize("kind", ["parallel", "fanflat"])
@pytest.mark.parametrize("opts", [
    {"FilterType": "ram-lak"}, {"FilterType": "none"},
    {"FilterType": "tukey", "FilterParameter": -1.0},
    {"FilterType": "gaussian", "FilterParameter": -1.0},
    {"FilterType": "blackman", "FilterParameter": -1.0},
    {"FilterType": "kaiser", "FilterParameter": -1.0},
    {"FilterType": "shepp-logan", "FilterD": 1.0}, {"FilterType": "cosine", "FilterD": 1.0},
    {"FilterType": "hamming", "FilterD": 1.0}, {"FilterType": "hann", "FilterD": 1.0},
], ids=lambda o: "-".join(str(v) for v in o.values()))
def test_2d_fbp_filter_options(kind, opts):
    out = _output(_config2("FBP_CUDA", _pg2(kind), opts))
    assert np.all(np.isfinite(out))
    assert not np.allclose(out, INIT)


def _custom_filter(kind, n_det, n_angles, rng):
    """A user filter as the reference's tests build one: a ramp over the
    half spectrum of the padded transform, or a real-space ramp kernel of
    n_det taps; 'sinogram' kinds scale one row per angle."""
    if kind in ("projection", "sinogram"):
        full = 1 << int(np.ceil(np.log2(2 * n_det)))
        row = np.linspace(0.0, 1.0, full // 2 + 1)
    else:
        row = np.zeros(n_det)
        c = n_det // 2
        row[c] = 0.5
        for k in range(1, n_det // 2, 2):
            if c + k < n_det:
                row[c + k] = row[c - k] = -2.0 / (np.pi * k) ** 2
    if kind in ("sinogram", "rsinogram"):
        return np.outer(rng.random(n_angles), row)
    return row[None, :]


@pytest.mark.parametrize("kind", ["parallel", "fanflat"])
@pytest.mark.parametrize("ftype", ["projection", "sinogram", "rprojection", "rsinogram"])
def test_2d_fbp_custom_filters(kind, ftype):
    a = _astra()
    k = _custom_filter(ftype, 40, 180, np.random.default_rng(3))
    fid = a.data2d.create("-sino", a.create_proj_geom("parallel", 1.0, k.shape[1],
                                                      np.zeros(k.shape[0])), k)
    try:
        out = _output(_config2("FBP_CUDA", _pg2(kind), {"FilterType": ftype, "FilterSinogramId": fid}))
    finally:
        a.data2d.delete(fid)
    assert np.all(np.isfinite(out))
    assert not np.allclose(out, INIT)


def test_2d_short_scan_changes_fbp():
    pg = _pg2("short_scan")
    plain = _output(_config2("FBP_CUDA", pg))
    short = _output(_config2("FBP_CUDA", pg, {"ShortScan": True}))
    assert not np.allclose(short, INIT)
    assert not np.allclose(short, plain)


@pytest.mark.parametrize("atype", ["SIRT_CUDA", "SART_CUDA"])
def test_2d_min_max_constraint_exact(atype):
    pg = _pg2("parallel")
    free = _output(_config2(atype, pg))
    boxed = _output(_config2(atype, pg, {"MinConstraint": 0.0, "MaxConstraint": 0.125}))
    assert free.min() < 0.0 and free.max() > 0.125
    assert boxed.min() == 0.0
    assert boxed.max() == 0.125


@pytest.mark.parametrize("atype", ["SIRT_CUDA", "SART_CUDA", "CGLS_CUDA"])
def test_2d_reconstruction_mask(atype):
    a = _astra()
    mask = np.random.default_rng(5).random((50, 60)) > 0.1
    mid = a.data2d.create("-vol", _vg2(), mask)
    try:
        out = _output(_config2(atype, _pg2("parallel"), {"ReconstructionMaskId": mid}))
        m = a.data2d.get(mid) > 0
    finally:
        a.data2d.delete(mid)
    assert not np.allclose(out, INIT)
    assert np.array_equal(out[~m], np.full((~m).sum(), INIT, np.float32))
    assert not np.allclose(out[m], INIT)


def test_2d_sinogram_mask():
    a = _astra()
    pg = _pg2("parallel")
    mid = a.data2d.create("-sino", pg, np.random.default_rng(6).random((180, 40)) > 0.1)
    try:
        plain = _output(_config2("SIRT_CUDA", pg))
        masked = _output(_config2("SIRT_CUDA", pg, {"SinogramMaskId": mid}))
    finally:
        a.data2d.delete(mid)
    assert not np.allclose(masked, INIT)
    assert not np.allclose(masked, plain)


@pytest.mark.parametrize("order", ["random", "sequential", "custom"])
def test_2d_sart_projection_order(order):
    opts = {"ProjectionOrder": order}
    if order == "custom":
        # 0, 5, 10, ..., 175, 1, 6, ...: every angle once, stride 5
        opts["ProjectionOrderList"] = np.concatenate([np.arange(r, 180, 5) for r in range(5)])
    out = _output(_config2("SART_CUDA", _pg2("parallel"), opts))
    assert np.all(np.isfinite(out))
    assert not np.allclose(out, INIT)


def test_2d_sart_full_sweep_order_independent_of_listing():
    """A 'custom' list equal to 0..179 is the 'sequential' order."""
    pg = _pg2("parallel")
    seq = _output(_config2("SART_CUDA", pg, {"ProjectionOrder": "sequential"}))
    lst = _output(_config2("SART_CUDA", pg, {"ProjectionOrder": "custom",
                                             "ProjectionOrderList": np.arange(180)}))
    assert np.array_equal(seq, lst)


def _residual(dim, pg, vg, x, p):
    """||p - FP(x)|| with the API's own FP on the same geometries."""
    a = _astra()
    d = a.data3d if dim == 3 else a.data2d
    vid = d.create("-vol", vg, x)
    sid = d.create("-sino", pg, 0.0)
    cfg = a.astra_dict("FP3D_CUDA" if dim == 3 else "FP_CUDA")
    cfg["VolumeDataId"], cfg["ProjectionDataId"] = vid, sid
    aid = a.algorithm.create(cfg)
    a.algorithm.run(aid)
    fx = d.get(sid).astype(np.float64)
    a.algorithm.delete(aid)
    d.delete([vid, sid])
    return float(np.linalg.norm(np.asarray(p, np.float64) - fx))


@pytest.mark.parametrize("atype", ["SIRT_CUDA", "SART_CUDA", "CGLS_CUDA", "EM_CUDA"])
def test_2d_get_res_norm(atype):
    a = _astra()
    pg = _pg2("parallel")
    cfg = _config2(atype, pg, {"ProjectionOrder": "sequential"} if atype == "SART_CUDA" else None)
    aid = a.algorithm.create(cfg)
    a.algorithm.run(aid, 2)
    r1 = a.algorithm.get_res_norm(aid)
    r2 = a.algorithm.get_res_norm(aid)
    x = a.data2d.get(cfg["ReconstructionDataId"])
    a.algorithm.delete(aid)
    ref = _residual(2, pg, _vg2(), x, _sino2_input())
    a.data2d.delete([cfg["ProjectionDataId"], cfg["ReconstructionDataId"]])
    assert r1 > 0.0 and r1 == r2
    assert abs(r1 - ref) <= 1e-4 * ref, (r1, ref)


@pytest.mark.parametrize("atype", ["SIRT_CUDA", "SART_CUDA", "CGLS_CUDA", "EM_CUDA"])
def test_2d_get_res_norm_leaves_state(atype):
    """run 2 + run 2 then a fresh run 2 gives the same image whether or not
    get_res_norm was called between the runs."""
    a = _astra()
    opts = {"ProjectionOrder": "sequential"} if atype == "SART_CUDA" else {}
    pg = _pg2("parallel")
    outs = []
    for probe in (False, True):
        cfg = _config2(atype, pg, opts)
        aid = a.algorithm.create(cfg)
        a.algorithm.run(aid, 2)
        if probe:
            r1 = a.algorithm.get_res_norm(aid)
            r2 = a.algorithm.get_res_norm(aid)
            assert r1 > 0.0 and np.isclose(r1, r2)
        a.algorithm.run(aid, 2)
        a.algorithm.delete(aid)
        outs.append(_output(cfg, iters=2))
    assert np.array_equal(outs[0], outs[1])


# --------------------------------------------------------------- 3D ------
def _vg3():
    a = _astra()
    return a.create_vol_geom(40, 30, 50, -15.0 + 1, 15.0 + 1, -20.0 + 2, 20.0 + 2,
                             -25.0 + 3, 25.0 + 3)


def _pg3(kind):
    a = _astra()
    if kind.startswith("parallel3d"):
        g = a.create_proj_geom("parallel3d", 1.0, 1.0, 20, 45, A3)
    elif kind in ("cone", "cone_vec", "cyl_cone_vec"):
        g = a.create_proj_geom("cone", 1.0, 1.0, 20, 45, A3, 100.0, 100.0)
    elif kind == "short_scan":
        half = np.arctan2(22.5, 200.0)
        g = a.create_proj_geom("cone", 1.0, 1.0, 20, 45, np.linspace(0, np.pi + 2 * half, 180),
                               100.0, 100.0)
    else:
        raise ValueError(kind)
    if kind == "cyl_cone_vec":
        v = a.geom_2vec(g)["Vectors"]
        return a.create_proj_geom("cyl_cone_vec", 20, 45,
                                  np.hstack([v, np.full((v.shape[0], 1), 200.0)]))
    return a.geom_2vec(g) if kind.endswith("_vec") else g


def _proj3_input():
    p = np.full((20, 180, 45), -10.0)
    p[5:15, :, 11:34] = 10.0
    return p


def _config3(atype, pg, opts=None):
    a = _astra()
    cfg = a.astra_dict(atype)
    vid = a.data3d.create("-vol", _vg3(), INIT)
    if atype.startswith("FP"):
        cfg["VolumeDataId"] = vid
        cfg["ProjectionDataId"] = a.data3d.create("-sino", pg, INIT)
    else:
        cfg["ReconstructionDataId"] = vid
        cfg["ProjectionDataId"] = a.data3d.create("-sino", pg, _proj3_input())
    if opts:
        cfg["option"] = dict(opts)
    return cfg


@pytest.mark.parametrize("kind", ["parallel3d", "parallel3d_vec", "cone", "cone_vec", "cyl_cone_vec"])
@pytest.mark.parametrize("atype", ["FP3D_CUDA", "BP3D_CUDA", "FDK_CUDA", "SIRT3D_CUDA", "CGLS3D_CUDA"])
def test_3d_algorithm_changes_output(kind, atype):
    if atype == "FDK_CUDA" and kind not in ("cone", "cone_vec"):
        a = _astra()
        cfg = _config3(atype, _pg3(kind))
        with pytest.raises(a.AstraError, match="cone"):
            a.algorithm.create(cfg)
        a.data3d.delete([cfg["ProjectionDataId"], cfg["ReconstructionDataId"]])
        return
    out = _output(_config3(atype, _pg3(kind)))
    assert out.shape == ((20, 180, 45) if atype == "FP3D_CUDA" else (50, 40, 30))
    assert np.all(np.isfinite(out))
    assert not np.allclose(out, INIT)


@pytest.mark.parametrize("kind", ["parallel3d", "cone"])
@pytest.mark.parametrize("atype", ["FP3D_CUDA", "SIRT3D_CUDA", "CGLS3D_CUDA"])
def test_3d_detector_supersampling(kind, atype):
    pg = _pg3(kind)
    plain = _output(_config3(atype, pg))
    ss = _output(_config3(atype, pg, {"DetectorSuperSampling": 3}))
    assert not np.allclose(ss, INIT)
    assert not np.allclose(ss, plain)


@pytest.mark.parametrize("kind", ["parallel3d", "cone"])
@pytest.mark.parametrize("atype", ["BP3D_CUDA", "FDK_CUDA", "SIRT3D_CUDA", "CGLS3D_CUDA"])
def test_3d_voxel_supersampling(kind, atype):
    if atype == "FDK_CUDA" and kind == "parallel3d":
        pytest.skip("FDK needs a cone geometry")
    pg = _pg3(kind)
    plain = _output(_config3(atype, pg))
    ss = _output(_config3(atype, pg, {"VoxelSuperSampling": 3}))
    assert not np.allclose(ss, INIT)
    if atype == "FDK_CUDA":
        # The reference's FDK job back projects with one ray per voxel
        # whatever the projector's supersampling
        # (src/CompositeGeometryManager.cpp:1237-1259 calls astraCUDA3d::FDK
        # without the projector's parameters), so the option leaves FDK as is.
        assert np.array_equal(ss, plain)
    else:
        assert not np.allclose(ss, plain)


@pytest.mark.parametrize("opts", [
    {"FilterType": "ram-lak"}, {"FilterType": "none"},
    {"FilterType": "tukey", "FilterParameter": -1.0},
    {"FilterType": "kaiser", "FilterParameter": -1.0},
    {"FilterType": "shepp-logan", "FilterD": 1.0}, {"FilterType": "hann", "FilterD": 1.0},
], ids=lambda o: "-".join(str(v) for v in o.values()))
def test_3d_fdk_filter_options(opts):
    out = _output(_config3("FDK_CUDA", _pg3("cone"), opts))
    assert np.all(np.isfinite(out))
    assert not np.allclose(out, INIT)


@pytest.mark.parametrize("ftype", ["projection", "sinogram", "rprojection", "rsinogram"])
def test_3d_fdk_custom_filters(ftype):
    a = _astra()
    k = _custom_filter(ftype, 45, 180, np.random.default_rng(4))
    fid = a.data2d.create("-sino", a.create_proj_geom("parallel", 1.0, k.shape[1],
                                                      np.zeros(k.shape[0])), k)
    try:
        out = _output(_config3("FDK_CUDA", _pg3("cone"), {"FilterType": ftype,
                                                           "FilterSinogramId": fid}))
    finally:
        a.data2d.delete(fid)
    assert np.all(np.isfinite(out))
    assert not np.allclose(out, INIT)


def test_3d_short_scan_changes_fdk():
    pg = _pg3("short_scan")
    plain = _output(_config3("FDK_CUDA", pg))
    short = _output(_config3("FDK_CUDA", pg, {"ShortScan": True}))
    assert not np.allclose(short, INIT)
    assert not np.allclose(short, plain)


def test_3d_min_max_constraint_exact():
    pg = _pg3("parallel3d")
    free = _output(_config3("SIRT3D_CUDA", pg))
    boxed = _output(_config3("SIRT3D_CUDA", pg, {"MinConstraint": 0.0, "MaxConstraint": 0.125}))
    assert free.min() < 0.0 and free.max() > 0.125
    assert boxed.min() == 0.0
    assert boxed.max() == 0.125


@pytest.mark.parametrize("atype", ["SIRT3D_CUDA", "CGLS3D_CUDA"])
def test_3d_reconstruction_mask(atype):
    a = _astra()
    mask = np.random.default_rng(7).random((50, 40, 30)) > 0.1
    mid = a.data3d.create("-vol", _vg3(), mask)
    try:
        out = _output(_config3(atype, _pg3("parallel3d"), {"ReconstructionMaskId": mid}))
        m = a.data3d.get(mid) > 0
    finally:
        a.data3d.delete(mid)
    assert not np.allclose(out, INIT)
    assert np.array_equal(out[~m], np.full((~m).sum(), INIT, np.float32))


def test_3d_sinogram_mask():
    a = _astra()
    pg = _pg3("parallel3d")
    mid = a.data3d.create("-sino", pg, np.random.default_rng(8).random((20, 180, 45)) > 0.1)
    try:
        plain = _output(_config3("SIRT3D_CUDA", pg))
        masked = _output(_config3("SIRT3D_CUDA", pg, {"SinogramMaskId": mid}))
    finally:
        a.data3d.delete(mid)
    assert not np.allclose(masked, INIT)
    assert not np.allclose(masked, plain)


@pytest.mark.parametrize("atype", ["SIRT3D_CUDA", "CGLS3D_CUDA"])
def test_3d_get_res_norm(atype):
    a = _astra()
    pg = _pg3("parallel3d")
    outs = []
    for probe in (False, True):
        cfg = _config3(atype, pg)
        aid = a.algorithm.create(cfg)
        a.algorithm.run(aid, 2)
        if probe:
            r1 = a.algorithm.get_res_norm(aid)
            r2 = a.algorithm.get_res_norm(aid)
            x = a.data3d.get(cfg["ReconstructionDataId"])
            ref = _residual(3, pg, _vg3(), x, _proj3_input())
            assert r1 > 0.0 and r1 == r2
            assert abs(r1 - ref) <= 1e-4 * ref, (r1, ref)
        a.algorithm.run(aid, 2)
        a.algorithm.delete(aid)
        outs.append(_output(cfg, iters=2))
    assert np.array_equal(outs[0], outs[1])
