"""Cone beam onto a cylindrical detector (cyl_cone_vec) on the GPU against
the FP64 oracle: FP in all three direction groups with anisotropic voxels
and a shifted window, BP, the matched adjoint, SIRT3D, and the kernels the
reference refuses for this geometry."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _vecs(na, sod, odd, R, du=1.2, dv=1.1, seed=None):
    rng = np.random.default_rng(seed) if seed is not None else None
    out = []
    for i, t in enumerate(np.linspace(0, 2 * np.pi, na, endpoint=False)):
        if rng is not None:
            d = rng.normal(size=3)
            if i % 3 == 2:
                d[2] = 3.0 * np.sign(d[2] + 1e-3)
            d /= np.linalg.norm(d)
            u = np.cross(d, [0.3, 0.2, 1.0]); u /= np.linalg.norm(u)
            v = np.cross(d, u)
            out.append(np.r_[d * sod, -d * odd, u * du, v * dv, R])
        else:
            c, s = np.cos(t), np.sin(t)
            out.append(np.r_[s * sod, -c * sod, 0.0, -s * odd, c * odd, 0.0, c * du, s * du, 0.0,
                             0.0, 0.0, dv, R])
    return np.array(out)


def _setup(vecs, nu, nv, vg_args):
    import astra
    from oracle import geometry as OG
    vg = astra.create_vol_geom(*vg_args)
    pg = astra.create_proj_geom("cyl_cone_vec", nv, nu, vecs)
    ov, vs = OG.translate_cylcone(vecs, vg)
    return astra, vg, pg, ov, vs


def _fp(astra, pid, vol, shape):
    import torch
    from astra import experimental
    out = torch.empty(shape, dtype=torch.float32, device="cuda")
    experimental.direct_FP3D(pid, torch.from_numpy(np.ascontiguousarray(vol, dtype=np.float32)).cuda(), out)
    return out.cpu().numpy()


def _bp(astra, pid, proj, shape):
    import torch
    from astra import experimental
    out = torch.empty(shape, dtype=torch.float32, device="cuda")
    experimental.direct_BP3D(pid, out, torch.from_numpy(np.ascontiguousarray(proj, dtype=np.float32)).cuda())
    return out.cpu().numpy()


def test_cylcone_fp_bp_circular():
    from astra import phantom
    from oracle import proj3d as O3
    n, na, nu, nv = 32, 36, 48, 40
    vecs = _vecs(na, 120.0, 80.0, 200.0)
    astra, vg, pg, ov, vs = _setup(vecs, nu, nv, (n, n, n))
    pid = astra.create_projector("cuda3d", pg, vg)
    vol = phantom.ellipsoids_3d(n, seed=91)
    got = _fp(astra, pid, vol, (nv, na, nu))
    assert rel_l2(got, O3.cylcone_fp(vol, ov, nu, nv, vs)) <= 1e-5
    proj = O3.cylcone_fp(vol, ov, nu, nv, vs).astype(np.float32)
    gotb = _bp(astra, pid, proj, (n, n, n))
    assert rel_l2(gotb, O3.cylcone_bp(proj, ov, (n, n, n), vs)) <= 1e-5


def test_cylcone_all_directions_anisotropic():
    from oracle import proj3d as O3
    na, nu, nv = 18, 30, 26
    vecs = _vecs(na, 90.0, 45.0, 140.0, seed=92)
    astra, vg, pg, ov, vs = _setup(vecs, nu, nv, (20, 24, 16, -15, 15, -10, 10, -6, 10))
    assert {O3.cylcone_direction(x) for x in ov} == {0, 1, 2}
    pid = astra.create_projector("cuda3d", pg, vg)
    rng = np.random.default_rng(93)
    vol = rng.random((16, 20, 24), dtype=np.float32)
    got = _fp(astra, pid, vol, (nv, na, nu))
    assert rel_l2(got, O3.cylcone_fp(vol, ov, nu, nv, vs)) <= 1e-5
    proj = O3.cylcone_fp(vol, ov, nu, nv, vs).astype(np.float32)
    gotb = _bp(astra, pid, proj, (16, 20, 24))
    assert rel_l2(gotb, O3.cylcone_bp(proj, ov, (16, 20, 24), vs)) <= 1e-5


def test_cylcone_matched_adjoint():
    import torch
    from astra import autograd as AG
    from oracle import proj3d as O3
    na, nu, nv = 18, 30, 26
    vecs = _vecs(na, 90.0, 45.0, 140.0, seed=94)
    astra, vg, pg, ov, vs = _setup(vecs, nu, nv, (20, 24, 16, -15, 15, -10, 10, -6, 10))
    pid = astra.create_projector("cuda3d", pg, vg)
    g = torch.Generator().manual_seed(95)
    x = torch.rand((16, 20, 24), generator=g).cuda()
    y = torch.rand((nv, na, nu), generator=g).cuda()
    aty = AG.back_project(y, pid)
    lhs = float((AG.forward_project(x, pid).double() * y.double()).sum())
    rhs = float((x.double() * aty.double()).sum())
    assert abs(lhs - rhs) / abs(lhs) <= 1e-5
    # the matched BP is the transpose of the oracle FP: <FP x, y> in FP64
    ref = float((O3.cylcone_fp(x.cpu().numpy().astype(np.float64), ov, nu, nv, vs) *
                 y.cpu().numpy()).sum())
    assert abs(ref - rhs) / abs(ref) <= 1e-5


def test_cylcone_sirt3d():
    from astra import phantom
    from oracle import algorithms as OA, proj3d as O3
    n, na, nu, nv = 24, 30, 36, 32
    vecs = _vecs(na, 100.0, 60.0, 160.0)
    astra, vg, pg, ov, vs = _setup(vecs, nu, nv, (n, n, n))
    proj = O3.cylcone_fp(phantom.ellipsoids_3d(n, seed=96), ov, nu, nv, vs).astype(np.float32)
    sid = astra.data3d.create("-sino", pg, proj)
    rid = astra.data3d.create("-vol", vg, 0)
    cfg = astra.astra_dict("SIRT3D_CUDA")
    cfg.update(ProjectionDataId=sid, ReconstructionDataId=rid)
    aid = astra.algorithm.create(cfg)
    astra.algorithm.run(aid, 10)
    ref = OA.sirt(lambda x: O3.cylcone_fp(x, ov, nu, nv, vs),
                  lambda y: O3.cylcone_bp(y, ov, (n, n, n), vs), proj, np.zeros((n, n, n)), 10)
    assert rel_l2(astra.data3d.get(rid), ref) <= 1e-4


def test_cylcone_refused_kernels():
    import astra
    vecs = _vecs(6, 100.0, 60.0, 160.0)
    vg = astra.create_vol_geom(8, 8, 8)
    pg = astra.create_proj_geom("cyl_cone_vec", 10, 12, vecs)
    sid = astra.data3d.create("-sino", pg, 0)
    rid = astra.data3d.create("-vol", vg, 0)
    cfg = astra.astra_dict("FDK_CUDA")
    cfg.update(ProjectionDataId=sid, ReconstructionDataId=rid)
    with pytest.raises(astra.AstraError):
        astra.algorithm.create(cfg)
    pid = astra.create_projector("cuda3d", pg, vg, {"DetectorSuperSampling": 2})
    cfg = astra.astra_dict("SIRT3D_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
    with pytest.raises(astra.AstraError):
        aid = astra.algorithm.create(cfg)
        astra.algorithm.run(aid, 1)
    pid2 = astra.create_projector("cuda3d", pg, vg, {"ProjectionKernel": "2d_weighting"})
    import torch
    from astra import experimental
    with pytest.raises(astra.AstraError):
        experimental.direct_BP3D(pid2, torch.zeros((8, 8, 8), device="cuda"),
                                 torch.zeros((10, 6, 12), device="cuda"))
