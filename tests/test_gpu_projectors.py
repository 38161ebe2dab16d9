"""GPU parity of the HIP projectors against the FP64 oracle.

Bar: relative L2 <= 1e-5 for one FP or BP.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _astra():
    import astra
    return astra


@pytest.fixture(scope="module")
def cone_case():
    from astra import phantom
    from oracle import geometry as OG

    astra = _astra()
    n, na, nd = 64, 90, 96
    angles = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, angles, 300.0, 150.0)
    ovec, vs = OG.normalise_3d(OG.gen_cone_projections(angles, nd, nd, 300.0, 150.0, 1.0, 1.0),
                               "cone", OG.vol_geom_3d(n, n, n))
    vol = phantom.ellipsoids_3d(n, seed=3)
    return dict(n=n, na=na, nd=nd, vg=vg, pg=pg, ovec=ovec, vs=vs, vol=vol)


def _fp3d(pid, vol, shape):
    import torch
    astra = _astra()
    vt = torch.from_numpy(vol).cuda()
    pt = torch.full(shape, 7.0, dtype=torch.float32, device="cuda")  # SET must overwrite
    astra.experimental.direct_FP3D(pid, vt, pt)
    return pt.cpu().numpy()


def _bp3d(pid, proj, shape):
    import torch
    astra = _astra()
    pt = torch.from_numpy(np.ascontiguousarray(proj, dtype=np.float32)).cuda()
    vt = torch.full(shape, 7.0, dtype=torch.float32, device="cuda")
    astra.experimental.direct_BP3D(pid, vt, pt)
    return vt.cpu().numpy()


def test_cone_fp(cone_case):
    from oracle import proj3d as O3
    c = cone_case
    astra = _astra()
    pid = astra.create_projector("cuda3d", c["pg"], c["vg"])
    got = _fp3d(pid, c["vol"], (c["nd"], c["na"], c["nd"]))
    ref = O3.cone_fp(c["vol"], c["ovec"], c["nd"], c["nd"], c["vs"])
    e = rel_l2(got, ref)
    print("cone FP rel-L2", e)
    assert e <= TOL


@pytest.mark.parametrize("kernel", ["default", "fdk", "2d"])
def test_cone_bp(cone_case, kernel):
    import torch
    from astra import _core, _lib
    from oracle import proj3d as O3
    c = cone_case
    astra = _astra()
    pid = astra.create_projector("cuda3d", c["pg"], c["vg"])
    rng = np.random.default_rng(4)
    proj = rng.random((c["nd"], c["na"], c["nd"]), dtype=np.float32)
    kmap = {"default": _lib.KER3D_DEFAULT, "fdk": _lib.KER3D_FDK_WEIGHTING, "2d": _lib.KER3D_2D_WEIGHTING}
    p = astra.projector3d.get(pid)
    pt = torch.from_numpy(proj).cuda()
    vt = torch.zeros(p.vol_shape, dtype=torch.float32, device="cuda")
    _core.bp(p, pt, vt, _lib.MODE_SET, kernel=kmap[kernel])
    got = vt.cpu().numpy()
    ref = O3.cone_bp(proj, c["ovec"], (c["n"],) * 3, c["vs"], kernel=kernel)
    e = rel_l2(got, ref)
    print("cone BP", kernel, "rel-L2", e)
    assert e <= TOL


def test_cone_fp_host_arrays_and_mode_add(cone_case):
    """numpy in/out through DLPack (host staging) and MODE_ADD accumulation."""
    from oracle import proj3d as O3
    c = cone_case
    astra = _astra()
    pid = astra.create_projector("cuda3d", c["pg"], c["vg"])
    out = np.ones((c["nd"], c["na"], c["nd"]), dtype=np.float32)
    astra.experimental.direct_FPBP3D(pid, c["vol"], out, astra.experimental.MODE_ADD, "FP")
    ref = O3.cone_fp(c["vol"], c["ovec"], c["nd"], c["nd"], c["vs"]) + 1.0
    assert rel_l2(out, ref) <= TOL


def test_par3d_fp_bp():
    from astra import phantom
    from oracle import geometry as OG, proj3d as O3
    astra = _astra()
    n, na, nu, nv = 48, 60, 72, 56
    angles = np.linspace(0, np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    pg = astra.create_proj_geom("parallel3d", 1.0, 1.0, nv, nu, angles)
    pid = astra.create_projector("cuda3d", pg, vg)
    ovec, vs = OG.normalise_3d(OG.gen_par3d_projections(angles, nu, nv, 1.0, 1.0), "par3d",
                               OG.vol_geom_3d(n, n, n))
    vol = phantom.ellipsoids_3d(n, seed=5)
    got = _fp3d(pid, vol, (nv, na, nu))
    ref = O3.par3d_fp(vol, ovec, nu, nv, vs)
    assert rel_l2(got, ref) <= TOL
    proj = np.random.default_rng(6).random((nv, na, nu), dtype=np.float32)
    gotb = _bp3d(pid, proj, (n, n, n))
    refb = O3.par3d_bp(proj, ovec, (n, n, n), vs)
    assert rel_l2(gotb, refb) <= TOL


def test_cone_vec_all_directions():
    """Tilted cone_vec geometry whose angles fall in all three principal-axis
    groups (x, y and z), a non-cubic volume and a shifted window."""
    from oracle import geometry as OG, proj3d as O3
    astra = _astra()
    rng = np.random.default_rng(7)
    na, nu, nv = 24, 40, 36
    vecs = np.zeros((na, 12))
    for i in range(na):
        d = rng.normal(size=3)
        if i % 3 == 2:
            d[2] = 3.0 * np.sign(d[2] + 1e-3)  # mostly along z
        d /= np.linalg.norm(d)
        src = d * 120.0
        det = -d * 60.0
        u = np.cross(d, [0.3, 0.2, 1.0]); u /= np.linalg.norm(u)
        v = np.cross(d, u)
        vecs[i] = np.concatenate([src, det, u * 1.1, v * 0.9])
    vg = astra.create_vol_geom(30, 34, 26, -17, 17, -14, 16, -13, 13)
    pg = astra.create_proj_geom("cone_vec", nv, nu, vecs)
    pid = astra.create_projector("cuda3d", pg, vg)
    ov = OG.conevec_from_astra(vecs, nu, nv)
    ovec, vs = OG.normalise_3d(ov, "cone", vg)
    dirs = {O3.cone_direction(ovec[a], nu, nv) for a in range(na)}
    assert dirs == {0, 1, 2}
    vol = rng.random((26, 30, 34), dtype=np.float32)
    got = _fp3d(pid, vol, (nv, na, nu))
    ref = O3.cone_fp(vol, ovec, nu, nv, vs)
    assert rel_l2(got, ref) <= TOL
    proj = rng.random((nv, na, nu), dtype=np.float32)
    gotb = _bp3d(pid, proj, (26, 30, 34))
    refb = O3.cone_bp(proj, ovec, (26, 30, 34), vs)
    assert rel_l2(gotb, refb) <= TOL


def test_noncube_voxels():
    from astra import phantom
    from oracle import geometry as OG, proj3d as O3
    astra = _astra()
    n, na, nd = 32, 30, 48
    angles = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n, -20, 20, -16, 16, -8, 8)  # voxel 1.25 x 1 x 0.5
    pg = astra.create_proj_geom("cone", 1.0, 0.7, nd, nd, angles, 200.0, 100.0)
    pid = astra.create_projector("cuda3d", pg, vg)
    ovec, vs = OG.normalise_3d(OG.gen_cone_projections(angles, nd, nd, 200.0, 100.0, 1.0, 0.7),
                               "cone", vg)
    vol = phantom.ellipsoids_3d(n, seed=8)
    got = _fp3d(pid, vol, (nd, na, nd))
    ref = O3.cone_fp(vol, ovec, nd, nd, vs)
    assert rel_l2(got, ref) <= TOL
    proj = np.random.default_rng(9).random((nd, na, nd), dtype=np.float32)
    gotb = _bp3d(pid, proj, (n, n, n))
    refb = O3.cone_bp(proj, ovec, (n, n, n), vs)
    assert rel_l2(gotb, refb) <= TOL


# ------------------------------------------------------------------- 2D ---
@pytest.fixture(scope="module")
def case2d():
    from astra import phantom
    n = 64
    return dict(n=n, vol=phantom.ellipses_2d(n, seed=11), nd=96, na=90)


def _fp2d(pid, vol, shape):
    import torch
    astra = _astra()
    vt = torch.from_numpy(vol).cuda()
    st = torch.full(shape, 3.0, dtype=torch.float32, device="cuda")
    astra.experimental.direct_FP2D(pid, vt, st)
    return st.cpu().numpy()


def _bp2d(pid, sino, shape):
    import torch
    astra = _astra()
    st = torch.from_numpy(np.ascontiguousarray(sino, dtype=np.float32)).cuda()
    vt = torch.full(shape, 3.0, dtype=torch.float32, device="cuda")
    astra.experimental.direct_BP2D(pid, vt, st)
    return vt.cpu().numpy()


@pytest.mark.parametrize("geom", ["parallel", "fanflat"])
def test_2d_fp_bp(case2d, geom):
    from oracle import geometry as OG, proj2d as O2
    astra = _astra()
    n, na, nd = case2d["n"], case2d["na"], case2d["nd"]
    vg = astra.create_vol_geom(n, n)
    if geom == "parallel":
        angles = np.linspace(0, np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("parallel", 0.8, nd, angles)
        ovec, osc = OG.normalise_2d(OG.gen_par_projections(angles, nd, 0.8), "par", vg)
        fp_ref = lambda v: O2.par_fp(v, ovec, nd, osc)
        bp_ref = lambda s: O2.par_bp(s, ovec, (n, n), osc)
    else:
        angles = np.linspace(0, 2 * np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("fanflat", 1.0, nd, angles, 200.0, 100.0)
        ovec, osc = OG.normalise_2d(OG.gen_fan_projections(angles, nd, 200.0, 100.0, 1.0), "fan", vg)
        fp_ref = lambda v: O2.fan_fp(v, ovec, nd, osc)
        bp_ref = lambda s: O2.fan_bp(s, ovec, (n, n), osc)
    pid = astra.create_projector("cuda", pg, vg)
    got = _fp2d(pid, case2d["vol"], (na, nd))
    e = rel_l2(got, fp_ref(case2d["vol"]))
    print(geom, "FP rel-L2", e)
    assert e <= TOL
    sino = np.random.default_rng(12).random((na, nd), dtype=np.float32)
    gotb = _bp2d(pid, sino, (n, n))
    e = rel_l2(gotb, bp_ref(sino))
    print(geom, "BP rel-L2", e)
    assert e <= TOL


def test_fan_fbp_weighted_bp(case2d):
    import torch
    from astra import _core, _lib
    from oracle import geometry as OG, proj2d as O2
    astra = _astra()
    n, na, nd = case2d["n"], case2d["na"], case2d["nd"]
    angles = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n)
    pg = astra.create_proj_geom("fanflat", 1.0, nd, angles, 200.0, 100.0)
    ovec, osc = OG.normalise_2d(OG.gen_fan_projections(angles, nd, 200.0, 100.0, 1.0), "fan", vg)
    pid = astra.create_projector("cuda", pg, vg)
    p = astra.projector.get(pid)
    sino = np.random.default_rng(13).random((na, nd), dtype=np.float32)
    vt = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    _core.bp(p, torch.from_numpy(sino).cuda(), vt, _lib.MODE_SET, fbp_weighted=True)
    ref = O2.fan_bp(sino, ovec, (n, n), osc, fbp_weighted=True)
    assert rel_l2(vt.cpu().numpy(), ref) <= TOL


def test_par_angle0_column_sums():
    """Known answer of astra.test() (python/astra/tests.py:27-48): at angle
    0 the parallel FP equals the column sums."""
    astra = _astra()
    rng = np.random.default_rng(14)
    vol = rng.random((32, 32), dtype=np.float32)
    vg = astra.create_vol_geom(32, 32)
    pg = astra.create_proj_geom("parallel", 1.0, 32, np.array([0.0]))
    pid = astra.create_projector("cuda", pg, vg)
    sid, sino = astra.create_sino(vol, pid)
    assert np.allclose(sino[0], vol.sum(axis=0), rtol=1e-6, atol=1e-5)
    vg3 = astra.create_vol_geom(16, 16, 16)
    pg3 = astra.create_proj_geom("parallel3d", 1.0, 1.0, 16, 16, np.array([0.0]))
    v3 = rng.random((16, 16, 16), dtype=np.float32)
    _, p3 = astra.create_sino3d_gpu(v3, pg3, vg3)
    assert np.allclose(p3[:, 0, :], v3.sum(axis=1), rtol=1e-6, atol=1e-5)
