"""DetectorSuperSampling / VoxelSuperSampling on the GPU against the FP64
oracle (2D par / fan, 3D cone / par3d, FP and BP), inside SIRT, and the
matched adjoint of a supersampled forward projector (dot-product test)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _geo2d(kind, n, na, nd):
    import astra
    from oracle import geometry as OG
    vg = astra.create_vol_geom(n, n)
    if kind == "parallel":
        ang = np.linspace(0, np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("parallel", 1.0, nd, ang)
        v, osc = OG.normalise_2d(OG.gen_par_projections(ang, nd, 1.0), "parallel", vg)
    else:
        ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("fanflat", 1.0, nd, ang, 3.0 * n, 1.5 * n)
        v, osc = OG.normalise_2d(OG.gen_fan_projections(ang, nd, 3.0 * n, 1.5 * n, 1.0), "fan", vg)
    return vg, pg, v, osc


@pytest.mark.parametrize("kind", ["parallel", "fanflat"])
def test_2d_supersampled_fp_bp(kind):
    import astra
    from astra import phantom
    from oracle import proj2d as O2
    n, na, nd = 40, 36, 60
    vg, pg, v, osc = _geo2d(kind, n, na, nd)
    fp = O2.par_fp if kind == "parallel" else O2.fan_fp
    bp = O2.par_bp if kind == "parallel" else O2.fan_bp
    vol = phantom.ellipses_2d(n, seed=51)
    pid = astra.create_projector("cuda", pg, vg, {"DetectorSuperSampling": 3, "VoxelSuperSampling": 2})
    _, got = astra.create_sino(vol, pid)
    assert rel_l2(got, fp(vol, v, nd, osc, rays=3)) <= 1e-5
    y = fp(vol, v, nd, osc).astype(np.float32)
    _, gb = astra.create_backprojection(y, pid)
    assert rel_l2(gb, bp(y, v, (n, n), osc, ss=2)) <= 1e-5


@pytest.mark.parametrize("kind", ["cone", "parallel3d"])
def test_3d_supersampled_fp_bp(kind):
    import torch
    import astra
    from astra import experimental, phantom
    from oracle import geometry as OG, proj3d as O3
    n, na, nd = 20, 24, 28
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    if kind == "cone":
        pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, ang, 80.0, 40.0)
        v, vs = OG.normalise_3d(OG.gen_cone_projections(ang, nd, nd, 80.0, 40.0, 1.0, 1.0), "cone",
                                OG.vol_geom_3d(n, n, n))
        fp, bp = O3.cone_fp, O3.cone_bp
    else:
        pg = astra.create_proj_geom("parallel3d", 1.0, 1.0, nd, nd, ang)
        v, vs = OG.normalise_3d(OG.gen_par3d_projections(ang, nd, nd, 1.0, 1.0), "parallel3d",
                                OG.vol_geom_3d(n, n, n))
        fp, bp = O3.par3d_fp, O3.par3d_bp
    pid = astra.create_projector("cuda3d", pg, vg, {"DetectorSuperSampling": 2, "VoxelSuperSampling": 2})
    vol = phantom.ellipsoids_3d(n, seed=52)
    proj = torch.empty((nd, na, nd), dtype=torch.float32, device="cuda")
    experimental.direct_FP3D(pid, torch.from_numpy(vol).cuda(), proj)
    assert rel_l2(proj.cpu().numpy(), fp(vol, v, nd, nd, vs, ss=2)) <= 1e-5
    y = fp(vol, v, nd, nd, vs).astype(np.float32)
    out = torch.empty((n, n, n), dtype=torch.float32, device="cuda")
    experimental.direct_BP3D(pid, out, torch.from_numpy(y).cuda())
    assert rel_l2(out.cpu().numpy(), bp(y, v, (n, n, n), vs, ss=2)) <= 1e-5


def test_sirt_with_supersampling():
    import astra
    from astra import phantom
    from oracle import algorithms as OA, proj2d as O2
    n, na, nd = 32, 30, 48
    vg, pg, v, osc = _geo2d("fanflat", n, na, nd)
    sino = O2.fan_fp(phantom.ellipses_2d(n, seed=53), v, nd, osc).astype(np.float32)
    pid = astra.create_projector("cuda", pg, vg, {"DetectorSuperSampling": 2, "VoxelSuperSampling": 2})
    sid = astra.data2d.create("-sino", pg, sino)
    rid = astra.data2d.create("-vol", vg, 0)
    cfg = astra.astra_dict("SIRT_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
    astra.algorithm.run(astra.algorithm.create(cfg), 10)
    ref = OA.sirt(lambda x: O2.fan_fp(x, v, nd, osc, rays=2), lambda y: O2.fan_bp(y, v, (n, n), osc, ss=2),
                  sino, np.zeros((n, n)), 10)
    assert rel_l2(astra.data2d.get(rid), ref) <= 1e-4


def test_matched_adjoint_of_supersampled_fp():
    import torch
    import astra
    from astra import autograd as AG
    ang = np.linspace(0, 2 * np.pi, 30, endpoint=False)
    vg = astra.create_vol_geom(24, 24, 24)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, 32, 32, ang, 100.0, 50.0)
    pid = astra.create_projector("cuda3d", pg, vg, {"DetectorSuperSampling": 2})
    g = torch.Generator().manual_seed(54)
    x = torch.rand((24, 24, 24), generator=g).cuda()
    y = torch.rand((32, 30, 32), generator=g).cuda()
    lhs = float((AG.forward_project(x, pid).double() * y.double()).sum())
    rhs = float((x.double() * AG.back_project(y, pid).double()).sum())
    assert abs(lhs - rhs) / abs(lhs) <= 1e-5
    vg2 = astra.create_vol_geom(40, 40)
    pg2 = astra.create_proj_geom("fanflat", 1.0, 60, ang, 120.0, 60.0)
    pid2 = astra.create_projector("cuda", pg2, vg2, {"DetectorSuperSampling": 3})
    x2 = torch.rand((40, 40), generator=g).cuda()
    y2 = torch.rand((30, 60), generator=g).cuda()
    lhs = float((AG.forward_project(x2, pid2).double() * y2.double()).sum())
    rhs = float((x2.double() * AG.back_project(y2, pid2).double()).sum())
    assert abs(lhs - rhs) / abs(lhs) <= 1e-5


def test_par3d_sum_square_weights():
    import torch
    import astra
    from astra import experimental
    from oracle import geometry as OG, proj3d as O3
    n, na, nd = 20, 18, 26
    ang = np.linspace(0, np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    pg = astra.create_proj_geom("parallel3d", 1.0, 1.0, nd, nd, ang)
    v, vs = OG.normalise_3d(OG.gen_par3d_projections(ang, nd, nd, 1.0, 1.0), "parallel3d",
                            OG.vol_geom_3d(n, n, n))
    pid = astra.create_projector("cuda3d", pg, vg, {"ProjectionKernel": "sum_square_weights"})
    proj = torch.empty((nd, na, nd), dtype=torch.float32, device="cuda")
    experimental.direct_FP3D(pid, torch.zeros((n, n, n), device="cuda"), proj)
    assert rel_l2(proj.cpu().numpy(), O3.par3d_fp_sumsq(v, nd, nd, (n, n, n), vs)) <= 1e-5
