"""Memory-bounded 3D jobs on host data (astra.composite: the reference's
CompositeGeometryManager splitting) and GPULink raw-pointer linking, on the
GPU: a split job equals the unsplit one, for FP, BP (default and matched
kernels), FDK, cone / par3d / cylindrical detectors."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture
def small_memory():
    import astra
    from astra import composite as C
    calls = []
    orig = C._sub_projector

    def spy(p, pg, vg):
        calls.append((pg["Vectors"].shape[0] if "Vectors" in pg else None, vg["GridSliceCount"],
                      pg["DetectorRowCount"]))
        return orig(p, pg, vg)

    C._sub_projector = spy
    try:
        yield astra, calls
    finally:
        C._sub_projector = orig
        astra.set_gpu_index(0, memory=0)


def _geoms(kind, n=40, na=48, nd=56):
    import astra
    vg = astra.create_vol_geom(n, n + 4, n - 6, -22, 22, -20, 20, -18, 16)
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    if kind == "cone":
        pg = astra.create_proj_geom("cone", 1.0, 1.1, nd, nd + 2, ang, 150.0, 80.0)
    elif kind == "parallel3d":
        pg = astra.create_proj_geom("parallel3d", 1.2, 1.0, nd, nd + 2, ang[: na // 2])
    else:
        v = []
        for t in ang:
            c, s = np.cos(t), np.sin(t)
            v.append([s * 150, -c * 150, 0, -s * 80, c * 80, 0, c * 1.1, s * 1.1, 0, 0, 0, 1.0, 230.0])
        pg = astra.create_proj_geom("cyl_cone_vec", nd, nd + 2, np.array(v))
    return vg, pg


@pytest.mark.parametrize("kind", ["cone", "parallel3d", "cylcone"])
def test_split_fp_bp_equal_unsplit(small_memory, kind):
    astra, calls = small_memory
    from astra import phantom
    vg, pg = _geoms(kind)
    vol = np.zeros((34, 40, 44), dtype=np.float32)   # [z][y][x] of the 44 x 40 x 34 grid
    vol[:, :, 2:42] = phantom.ellipsoids_3d(40, seed=101)[:34]
    _, ref = astra.create_sino3d_gpu(vol, pg, vg)
    _, refb = astra.create_backprojection3d_gpu(ref, pg, vg)
    astra.set_gpu_index(0, memory=1024 * 1024)
    _, got = astra.create_sino3d_gpu(vol, pg, vg)
    n_fp = len(calls)
    _, gotb = astra.create_backprojection3d_gpu(ref, pg, vg)
    assert n_fp >= 2 and len(calls) >= n_fp + 2  # both jobs were split
    assert rel_l2(got, ref) <= 1e-6
    assert rel_l2(gotb, refb) <= 1e-6


@pytest.mark.parametrize("kind", ["cone", "parallel3d", "cylcone"])
def test_split_jobs_reduce_to_reachable_rows_and_slices(small_memory, kind):
    """A volume taller than the detector's view: the split FP skips the
    slices that project off the detector and the split BP uploads only the
    rows each slab reads; both still equal the unsplit jobs."""
    astra, calls = small_memory
    from astra import phantom
    import torch
    from astra import composite as C
    n, nz, na, nu, nv = 32, 128, 40, 48, 24
    vg = astra.create_vol_geom(n, n, nz)
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    if kind == "cone":
        pg = astra.create_proj_geom("cone", 1.0, 1.0, nv, nu, ang, 120.0, 60.0)
    elif kind == "parallel3d":
        pg = astra.create_proj_geom("parallel3d", 1.0, 1.0, nv, nu, ang)
    else:
        base = astra.geom_2vec(astra.create_proj_geom("cone", 1.0, 1.0, nv, nu, ang, 120.0, 60.0))
        pg = astra.create_proj_geom("cyl_cone_vec", nv, nu,
                                    np.hstack([base["Vectors"], np.full((na, 1), 170.0)]))
    vol = np.concatenate([phantom.ellipsoids_3d(n, seed=104 + k) for k in range(nz // n)])
    proj = np.random.default_rng(105).random((nv, na, nu), dtype=np.float32)
    _, ref = astra.create_sino3d_gpu(vol, pg, vg)
    _, refb = astra.create_backprojection3d_gpu(proj, pg, vg)
    astra.set_gpu_index(0, memory=1024 * 1024)
    _, got = astra.create_sino3d_gpu(vol, pg, vg)
    fp_calls = list(calls)
    _, gotb = astra.create_backprojection3d_gpu(proj, pg, vg)
    bp_calls = calls[len(fp_calls):]
    assert len(fp_calls) >= 2 and len(bp_calls) >= 2
    plan = C.plan_fp((nz, n, n), (nv, na, nu), C.budget(torch.device("cuda", 0)))
    assert sum(z for _, z, _ in fp_calls) < len(plan) * nz  # slices off the detector skipped
    pairs = sum(len(ch) for _, ch in C.plan_bp((nz, n, n), (nv, na, nu), C.budget(torch.device("cuda", 0))))
    # slabs read a subset of the rows, or none
    assert min(r for _, _, r in bp_calls) < nv or len(bp_calls) < pairs
    assert rel_l2(got, ref) <= 1e-6
    assert rel_l2(gotb, refb) <= 1e-6


def test_split_matched_bp_and_mode_add(small_memory):
    astra, calls = small_memory
    from astra import experimental
    vg, pg = _geoms("cone")
    rng = np.random.default_rng(102)
    proj = rng.random((56, 48, 58), dtype=np.float32)
    pid = astra.create_projector("cuda3d", pg, vg, {"ProjectionKernel": "matched_bp"})
    base = rng.random((34, 40, 44), dtype=np.float32)
    ref = base.copy()
    experimental.direct_FPBP3D(pid, ref, proj.copy(), experimental.MODE_ADD, "BP")
    astra.set_gpu_index(0, memory=1024 * 1024)
    got = base.copy()
    experimental.direct_FPBP3D(pid, got, proj.copy(), experimental.MODE_ADD, "BP")
    assert len(calls) > 2
    assert rel_l2(got, ref) <= 1e-6


def test_split_fdk(small_memory):
    astra, calls = small_memory
    from astra import phantom
    from oracle import fdk as OF, geometry as OG, proj3d as O3
    n, na, nd = 32, 40, 48
    vg = astra.create_vol_geom(n, n, n)
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, ang, 150.0, 75.0)
    ov, vs = OG.normalise_3d(OG.gen_cone_projections(ang, nd, nd, 150.0, 75.0, 1.0, 1.0), "cone",
                             OG.vol_geom_3d(n, n, n))
    proj = O3.cone_fp(phantom.ellipsoids_3d(n, seed=103), ov, nd, nd, vs).astype(np.float32)
    fixed = proj.nbytes + na * (nd + 4) * (nd + 4) * 8
    astra.set_gpu_index(0, memory=int((fixed + n * n * 4 * 10) / 0.9))
    sid = astra.data3d.create("-sino", pg, proj)
    rid = astra.data3d.create("-vol", vg, 0)
    cfg = astra.astra_dict("FDK_CUDA")
    cfg.update(ProjectionDataId=sid, ReconstructionDataId=rid)
    astra.algorithm.run(astra.algorithm.create(cfg))
    assert len(calls) >= 3
    assert rel_l2(astra.data3d.get(rid), OF.fdk(proj, ov, (n, n, n), vs)) <= 1e-5
    assert np.array_equal(astra.data3d.get(sid), proj)  # input untouched


def test_gpulink_pitched_volume_and_projections():
    import torch
    import astra
    from astra import phantom
    vg, pg = _geoms("cone")
    vol = np.ascontiguousarray(phantom.ellipsoids_3d(44, seed=104)[:34, :40, :44])
    _, ref = astra.create_sino3d_gpu(vol, pg, vg)
    vbuf = torch.zeros((34, 40, 48), device="cuda")       # rows padded to 48 floats
    vbuf[:, :, :44] = torch.from_numpy(vol).cuda()
    pbuf = torch.full((56, 48, 64), 7.0, device="cuda")   # rows padded to 64 floats
    vid = astra.data3d.link("-vol", vg, astra.GPULink(vbuf.data_ptr(), 44, 40, 34, 48 * 4))
    sid = astra.data3d.link("-sino", pg, astra.GPULink(pbuf.data_ptr(), 58, 48, 56, 64 * 4))
    cfg = astra.astra_dict("FP3D_CUDA")
    cfg.update(ProjectionDataId=sid, VolumeDataId=vid)
    astra.algorithm.run(astra.algorithm.create(cfg))
    torch.cuda.synchronize()
    assert rel_l2(pbuf[:, :, :58].cpu().numpy(), ref) <= 1e-6
    assert float(pbuf[:, :, 58:].min()) == 7.0 and float(pbuf[:, :, 58:].max()) == 7.0
    assert rel_l2(astra.data3d.get(sid), ref) <= 1e-6
    with pytest.raises(ValueError):
        astra.data3d.link("-vol", vg, astra.GPULink(vbuf.data_ptr(), 44, 40, 33, 48 * 4))
    astra.data3d.delete([vid, sid])


def test_pitched_entry_points_equal_contiguous():
    """astra_amd_{fp,bp}{2,3}d_pitched on padded rows equal the contiguous
    calls, MODE_ADD included, and leave the padding columns untouched."""
    import ctypes
    import torch
    import astra
    from astra import _core, _lib
    vg, pg = _geoms("cone")
    p3 = _core.Projector("cuda3d", pg, vg)
    dev = torch.device("cuda", 0)
    h3 = ctypes.c_void_p(p3.handle(dev))
    s = _core.stream_ptr(dev)
    rng = np.random.default_rng(105)
    vol = torch.from_numpy(rng.random((34, 40, 44), dtype=np.float32)).cuda()
    prm = p3.params3d()
    ref = torch.empty((56, 48, 58), device="cuda")
    _lib.call("astra_amd_fp3d", h3, _core.ptr(vol), _core.ptr(ref), ctypes.byref(prm), _lib.MODE_SET, s)
    vpad = torch.full((34, 40, 50), 3.0, device="cuda")
    vpad[:, :, :44] = vol
    ppad = torch.full((56, 48, 61), 5.0, device="cuda")
    _lib.call("astra_amd_fp3d_pitched", h3, _core.ptr(vpad), 50, _core.ptr(ppad), 61,
              ctypes.byref(prm), _lib.MODE_SET, s)
    torch.cuda.synchronize()
    assert torch.equal(ppad[:, :, :58], ref) and bool((ppad[:, :, 58:] == 5.0).all())
    refb = torch.ones((34, 40, 44), device="cuda")
    _lib.call("astra_amd_bp3d", h3, _core.ptr(ref), _core.ptr(refb), ctypes.byref(prm), _lib.MODE_ADD, s)
    vout = torch.full((34, 40, 50), 1.0, device="cuda")
    _lib.call("astra_amd_bp3d_pitched", h3, _core.ptr(ppad), 61, _core.ptr(vout), 50,
              ctypes.byref(prm), _lib.MODE_ADD, s)
    torch.cuda.synchronize()
    assert torch.equal(vout[:, :, :44], refb) and bool((vout[:, :, 44:] == 1.0).all())
    # 2D
    vg2 = astra.create_vol_geom(30, 26)
    pg2 = astra.create_proj_geom("fanflat", 1.0, 40, np.linspace(0, 6.2, 33), 80.0, 40.0)
    p2 = _core.Projector("cuda", pg2, vg2)
    h2 = ctypes.c_void_p(p2.handle(dev))
    prm2 = p2.params2d()
    img = torch.from_numpy(rng.random((30, 26), dtype=np.float32)).cuda()
    sino = torch.empty((33, 40), device="cuda")
    _lib.call("astra_amd_fp2d", h2, _core.ptr(img), _core.ptr(sino), ctypes.byref(prm2), _lib.MODE_SET, s)
    ipad = torch.zeros((30, 32), device="cuda")
    ipad[:, :26] = img
    spad = torch.full((33, 45), 2.0, device="cuda")
    _lib.call("astra_amd_fp2d_pitched", h2, _core.ptr(ipad), 32, _core.ptr(spad), 45,
              ctypes.byref(prm2), _lib.MODE_SET, s)
    bref = torch.empty((30, 26), device="cuda")
    _lib.call("astra_amd_bp2d", h2, _core.ptr(sino), _core.ptr(bref), ctypes.byref(prm2), _lib.MODE_SET, s)
    bpad = torch.full((30, 32), 4.0, device="cuda")
    _lib.call("astra_amd_bp2d_pitched", h2, _core.ptr(spad), 45, _core.ptr(bpad), 32,
              ctypes.byref(prm2), _lib.MODE_SET, s)
    torch.cuda.synchronize()
    assert torch.equal(spad[:, :40], sino) and bool((spad[:, 40:] == 2.0).all())
    assert torch.equal(bpad[:, :26], bref) and bool((bpad[:, 26:] == 4.0).all())
    with pytest.raises(astra.AstraError):
        _lib.call("astra_amd_fp2d_pitched", h2, _core.ptr(ipad), 20, _core.ptr(spad), 45,
                  ctypes.byref(prm2), _lib.MODE_SET, s)
