"""FBP / FDK with every filter of the reference (FilterType windows with
FilterD / FilterParameter, 'none', and the user filters projection /
sinogram / rprojection / rsinogram given by FilterSinogramId) against the
FP64 oracle.  Filter inputs follow the reference's own tests
(tests/python/unit/test_algorithms_2d.py:76-110, 271-309)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

N, NA, ND = 64, 90, 96


@pytest.fixture(scope="module")
def par():
    import astra
    from oracle import geometry as OG
    angles = np.linspace(0, np.pi, NA, endpoint=False)
    vg = astra.create_vol_geom(N, N)
    pg = astra.create_proj_geom("parallel", 1.0, ND, angles)
    ovec, osc = OG.normalise_2d(OG.gen_par_projections(angles, ND, 1.0), "par", vg)
    # sinogram data (projections of a phantom): windows that keep only high
    # frequencies turn white noise into a test of fp32 coordinate rounding
    from astra import phantom
    from oracle import proj2d as O2
    sino = O2.par_fp(phantom.ellipses_2d(N, seed=31), ovec, ND, osc).astype(np.float32)
    return dict(astra=astra, vg=vg, pg=pg, ovec=ovec, osc=osc, sino=sino)


def _fbp(c, **opts):
    astra = c["astra"]
    pid = astra.create_projector("cuda", c["pg"], c["vg"])
    sid = astra.data2d.create("-sino", c["pg"], c["sino"])
    rid = astra.data2d.create("-vol", c["vg"], 0)
    cfg = astra.astra_dict("FBP_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid, option=opts)
    aid = astra.algorithm.create(cfg)
    astra.algorithm.run(aid)
    out = astra.data2d.get(rid)
    astra.algorithm.delete(aid)
    astra.data2d.delete([sid, rid])
    astra.projector.delete(pid)
    return out


def _fbp_ref(c, **kw):
    from oracle import fdk as OF
    return OF.fbp_par(c["sino"], c["ovec"], (N, N), c["osc"], 1.0, **kw)


FILTERS = ["ram-lak", "shepp-logan", "cosine", "hamming", "hann", "tukey", "lanczos", "triangular",
           "gaussian", "barlett-hann", "blackman", "nuttall", "blackman-harris", "blackman-nuttall",
           "flat-top", "kaiser", "parzen"]


@pytest.mark.parametrize("name", FILTERS)
def test_fbp_windows(par, name):
    from oracle import fdk as OF
    got = _fbp(par, FilterType=name)
    ref = _fbp_ref(par, H=OF.windowed_filter(name, ND))
    assert rel_l2(got, ref) <= 1e-5


def test_fbp_filter_d_and_parameter(par):
    from oracle import fdk as OF
    for name, d, p in (("hann", 0.6, -1.0), ("shepp-logan", 0.8, -1.0), ("gaussian", 1.0, 0.25),
                       ("tukey", 1.0, 0.3), ("blackman", 1.0, 0.2), ("kaiser", 1.0, 2.0)):
        got = _fbp(par, FilterType=name, FilterD=d, FilterParameter=p)
        ref = _fbp_ref(par, H=OF.windowed_filter(name, ND, d, p))
        assert rel_l2(got, ref) <= 1e-5, name


def test_fbp_no_filter(par):
    got = _fbp(par, FilterType="none")
    assert rel_l2(got, _fbp_ref(par, filt="none")) <= 1e-5


def _fourier_filter():
    n = 1
    while n < 2 * ND:
        n *= 2
    return np.linspace(0, 1, n // 2 + 1).reshape(1, -1)


def _real_filter():
    k = np.zeros((1, ND))
    for i in range(ND // 4):
        f = np.pi * (2 * i + 1)
        k[0, ND // 2 + 2 * i + 1] = k[0, ND // 2 - (2 * i + 1)] = -2.0 / (f * f)
    k[0, ND // 2] = 0.5
    return k


@pytest.mark.parametrize("kind", ["projection", "sinogram", "rprojection", "rsinogram"])
def test_fbp_custom_filters(par, kind):
    from oracle import fdk as OF
    astra = par["astra"]
    base = _fourier_filter() if kind in ("projection", "sinogram") else _real_filter()
    if kind in ("sinogram", "rsinogram"):
        base = np.outer(np.random.default_rng(32).random(NA), base)
    fg = astra.create_proj_geom("parallel", 1.0, base.shape[1], np.zeros(base.shape[0]))
    fid = astra.data2d.create("-sino", fg, base.astype(np.float32))
    got = _fbp(par, FilterType=kind, FilterSinogramId=fid)
    H = OF.custom_filter(base.astype(np.float32), ND, kind.startswith("r"))
    assert rel_l2(got, _fbp_ref(par, H=H)) <= 1e-5
    astra.data2d.delete(fid)


def test_fan_fbp_window():
    import astra
    from oracle import fdk as OF, geometry as OG, proj2d as O2
    n, na, nd = 64, 90, 96
    angles = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n)
    pg = astra.create_proj_geom("fanflat", 1.0, nd, angles, 200.0, 100.0)
    ovec, osc = OG.normalise_2d(OG.gen_fan_projections(angles, nd, 200.0, 100.0, 1.0), "fan", vg)
    from astra import phantom
    sino = O2.fan_fp(phantom.ellipses_2d(n, seed=33), ovec, nd, osc).astype(np.float32)
    pid = astra.create_projector("cuda", pg, vg)
    sid = astra.data2d.create("-sino", pg, sino)
    rid = astra.data2d.create("-vol", vg, 0)
    cfg = astra.astra_dict("FBP_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid,
               option={"FilterType": "hamming"})
    astra.algorithm.run(astra.algorithm.create(cfg))
    ref = OF.fbp_fan(sino, ovec, (n, n), osc, 1.0, H=OF.windowed_filter("hamming", nd))
    assert rel_l2(astra.data2d.get(rid), ref) <= 1e-5


@pytest.mark.parametrize("kind", ["shepp-logan", "rsinogram"])
def test_fdk_filters(kind):
    import astra
    from oracle import fdk as OF, geometry as OG, proj3d as O3
    from astra import phantom
    n, na, nd = 32, 45, 48
    angles = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, angles, 150.0, 75.0)
    ovec, vs = OG.normalise_3d(OG.gen_cone_projections(angles, nd, nd, 150.0, 75.0, 1.0, 1.0),
                               "cone", OG.vol_geom_3d(n, n, n))
    proj = O3.cone_fp(phantom.ellipsoids_3d(n, seed=34), ovec, nd, nd, vs).astype(np.float32)
    opts = {"FilterType": kind}
    if kind == "rsinogram":
        base = np.outer(np.random.default_rng(35).random(na) + 0.5,
                        np.array([-0.1, -0.2, 1.0, -0.2, -0.1]))
        fg = astra.create_proj_geom("parallel", 1.0, base.shape[1], np.zeros(na))
        opts["FilterSinogramId"] = astra.data2d.create("-sino", fg, base.astype(np.float32))
        H = OF.custom_filter(base.astype(np.float32), nd, True)
    else:
        H = OF.windowed_filter(kind, nd)
    pid = astra.create_projector("cuda3d", pg, vg)
    sid = astra.data3d.create("-sino", pg, proj)
    rid = astra.data3d.create("-vol", vg, 0)
    cfg = astra.astra_dict("FDK_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid, option=opts)
    astra.algorithm.run(astra.algorithm.create(cfg))
    ref = OF.fdk(proj, ovec, (n, n, n), vs, H=H)
    assert rel_l2(astra.data3d.get(rid), ref) <= 1e-5


def test_fdk_short_scan():
    """FDK over pi + 2 x fan angle with Parker weights (ShortScan)."""
    import math
    import astra
    from astra import phantom
    from oracle import fdk as OF, geometry as OG, proj3d as O3
    n, nd, sod, odd = 32, 48, 150.0, 75.0
    fan = math.atan(nd / 2 / (sod + odd))
    angles = np.linspace(0, math.pi + 2 * fan, 60, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, angles, sod, odd)
    ovec, vs = OG.normalise_3d(OG.gen_cone_projections(angles, nd, nd, sod, odd, 1.0, 1.0),
                               "cone", OG.vol_geom_3d(n, n, n))
    proj = O3.cone_fp(phantom.ellipsoids_3d(n, seed=36), ovec, nd, nd, vs).astype(np.float32)
    pid = astra.create_projector("cuda3d", pg, vg)
    sid = astra.data3d.create("-sino", pg, proj)
    rid = astra.data3d.create("-vol", vg, 0)
    cfg = astra.astra_dict("FDK_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid,
               option={"ShortScan": True})
    astra.algorithm.run(astra.algorithm.create(cfg))
    ref = OF.fdk(proj, ovec, (n, n, n), vs, short_scan=True)
    assert rel_l2(astra.data3d.get(rid), ref) <= 1e-5


def test_fan_fbp_short_scan():
    import math
    import astra
    from astra import phantom
    from oracle import fdk as OF, geometry as OG, proj2d as O2
    n, nd, sod, odd = 64, 96, 200.0, 100.0
    fan = math.atan(nd / 2 / (sod + odd))
    angles = np.linspace(0, math.pi + 2 * fan, 120, endpoint=False)
    vg = astra.create_vol_geom(n, n)
    pg = astra.create_proj_geom("fanflat", 1.0, nd, angles, sod, odd)
    ovec, osc = OG.normalise_2d(OG.gen_fan_projections(angles, nd, sod, odd, 1.0), "fan", vg)
    sino = O2.fan_fp(phantom.ellipses_2d(n, seed=37), ovec, nd, osc).astype(np.float32)
    pid = astra.create_projector("cuda", pg, vg)
    sid = astra.data2d.create("-sino", pg, sino)
    rid = astra.data2d.create("-vol", vg, 0)
    cfg = astra.astra_dict("FBP_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid,
               option={"ShortScan": True})
    astra.algorithm.run(astra.algorithm.create(cfg))
    ref = OF.fbp_fan(sino, ovec, (n, n), osc, 1.0, short_scan=True)
    assert rel_l2(astra.data2d.get(rid), ref) <= 1e-5


def test_fbp_wide_detector_large_fft():
    """5000 detector bins: the row transform has 16384 points (128 KiB of
    LDS per workgroup, above the default 64 KiB).  Padded detector
    coordinates reach 5000, where an fp32 step is 4.9e-4 texel, so the BP
    of the filtered (high-pass) data is held to 5e-5 (the 1e-5 cases use
    detectors of a few hundred bins)."""
    import astra
    from oracle import fdk as OF, geometry as OG, proj2d as O2
    n, na, nd = 64, 24, 5000
    angles = np.linspace(0, np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n)
    pg = astra.create_proj_geom("parallel", 0.02, nd, angles)
    ovec, osc = OG.normalise_2d(OG.gen_par_projections(angles, nd, 0.02), "par", vg)
    from astra import phantom
    sino = O2.par_fp(phantom.ellipses_2d(n, seed=38), ovec, nd, osc).astype(np.float32)
    pid = astra.create_projector("cuda", pg, vg)
    sid = astra.data2d.create("-sino", pg, sino)
    rid = astra.data2d.create("-vol", vg, 0)
    cfg = astra.astra_dict("FBP_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
    astra.algorithm.run(astra.algorithm.create(cfg))
    ref = OF.fbp_par(sino, ovec, (n, n), osc, 1.0)
    assert rel_l2(astra.data2d.get(rid), ref) <= 5e-5


@pytest.fixture
def single_row():
    """Forces the single-row filter path (a real transform of length N as
    a complex one of length N/2; taken by itself above 16384 points) for
    the calls inside the test."""
    import os
    key = "ASTRA_AMD_FILTER_SINGLE_ROW"
    old = os.environ.get(key)
    os.environ[key] = "1"
    yield
    if old is None:
        os.environ.pop(key, None)
    else:
        os.environ[key] = old


@pytest.mark.parametrize("name", ["ram-lak", "hann"])
def test_single_row_path_fbp(par, single_row, name):
    from oracle import fdk as OF
    got = _fbp(par, FilterType=name)
    assert rel_l2(got, _fbp_ref(par, H=OF.windowed_filter(name, ND))) <= 1e-5


@pytest.mark.parametrize("kind", ["sinogram", "rsinogram"])
def test_single_row_path_custom_filters(par, single_row, kind):
    """Per-row (complex, Hermitian) spectra through the single-row path."""
    from oracle import fdk as OF
    astra = par["astra"]
    base = _fourier_filter() if kind == "sinogram" else _real_filter()
    base = np.outer(np.random.default_rng(36).random(NA), base)
    fg = astra.create_proj_geom("parallel", 1.0, base.shape[1], np.zeros(base.shape[0]))
    fid = astra.data2d.create("-sino", fg, base.astype(np.float32))
    got = _fbp(par, FilterType=kind, FilterSinogramId=fid)
    H = OF.custom_filter(base.astype(np.float32), ND, kind.startswith("r"))
    assert rel_l2(got, _fbp_ref(par, H=H)) <= 1e-5
    astra.data2d.delete(fid)


def test_single_row_path_fdk_weights(single_row):
    """The FDK pre-weights (and Parker weights) applied while loading a
    single row."""
    test_fdk_short_scan()


@pytest.fixture
def lds_pair_kernel():
    """Forces the LDS row-pair kernel for real spectra (the path of
    per-row spectra and of transforms above 4096 points) for the calls
    inside the test."""
    import os
    key = "ASTRA_AMD_FILTER_LDS_RADIX4"
    old = os.environ.get(key)
    os.environ[key] = "1"
    yield
    if old is None:
        os.environ.pop(key, None)
    else:
        os.environ[key] = old


def test_lds_pair_path_fdk_weights(lds_pair_kernel):
    """The FDK pre-weights and Parker weights applied while loading a row
    pair into the LDS kernel."""
    test_fdk_short_scan()


def test_fbp_detector_wider_than_a_row_pair():
    """10000 detector bins: the transform has 32768 points, whose row pair
    (256 KiB) exceeds the CU's LDS, so each row is filtered alone (128 KiB).
    Padded detector coordinates reach 10000, where an fp32 step is 9.8e-4
    texel, so the BP of the filtered data is held to 1e-4."""
    import astra
    from oracle import fdk as OF, geometry as OG, proj2d as O2
    from astra import phantom
    n, na, nd = 48, 12, 10000
    angles = np.linspace(0, np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n)
    pg = astra.create_proj_geom("parallel", 0.008, nd, angles)
    ovec, osc = OG.normalise_2d(OG.gen_par_projections(angles, nd, 0.008), "par", vg)
    sino = O2.par_fp(phantom.ellipses_2d(n, seed=39), ovec, nd, osc).astype(np.float32)
    pid = astra.create_projector("cuda", pg, vg)
    sid = astra.data2d.create("-sino", pg, sino)
    rid = astra.data2d.create("-vol", vg, 0)
    cfg = astra.astra_dict("FBP_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
    astra.algorithm.run(astra.algorithm.create(cfg))
    ref = OF.fbp_par(sino, ovec, (n, n), osc, 1.0)
    e = rel_l2(astra.data2d.get(rid), ref)
    print("FBP at 10000 bins rel-L2", e)
    assert e <= 1e-4
