"""Edge shapes on the GPU against the FP64 oracle: single slices, single
detector rows and columns, a single angle, one-voxel volumes, odd sizes,
and the filtered algorithms at one detector."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

SHAPES_3D = [
    # (nx, ny, nz, na, nu, nv)
    (9, 7, 1, 12, 11, 3),     # one z slice
    (5, 6, 7, 1, 8, 9),       # one angle
    (6, 6, 6, 10, 1, 7),      # one detector column
    (6, 6, 6, 10, 9, 1),      # one detector row
    (1, 1, 1, 5, 3, 3),       # one voxel
    (17, 5, 9, 23, 13, 11),   # odd everything
]


@pytest.mark.parametrize("kind", ["cone", "parallel3d"])
@pytest.mark.parametrize("shape", SHAPES_3D)
def test_3d_edge_shapes(kind, shape):
    import torch
    import astra
    from astra import experimental
    from oracle import geometry as OG, proj3d as O3
    nx, ny, nz, na, nu, nv = shape
    ang = np.linspace(0.3, 0.3 + 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(ny, nx, nz)
    if kind == "cone":
        pg = astra.create_proj_geom("cone", 1.0, 1.0, nv, nu, ang, 40.0, 20.0)
        ov, vs = OG.normalise_3d(OG.gen_cone_projections(ang, nu, nv, 40.0, 20.0, 1.0, 1.0), "cone", vg)
        fp, bp = O3.cone_fp, O3.cone_bp
    else:
        pg = astra.create_proj_geom("parallel3d", 1.0, 1.0, nv, nu, ang)
        ov, vs = OG.normalise_3d(OG.gen_par3d_projections(ang, nu, nv, 1.0, 1.0), "parallel3d", vg)
        fp, bp = O3.par3d_fp, O3.par3d_bp
    pid = astra.create_projector("cuda3d", pg, vg)
    rng = np.random.default_rng(sum(shape))
    vol = rng.random((nz, ny, nx), dtype=np.float32)
    proj = torch.empty((nv, na, nu), dtype=torch.float32, device="cuda")
    experimental.direct_FP3D(pid, torch.from_numpy(vol).cuda(), proj)
    ref = fp(vol, ov, nu, nv, vs)
    if np.linalg.norm(ref) > 0:
        assert rel_l2(proj.cpu().numpy(), ref) <= 1e-5
    y = rng.random((nv, na, nu), dtype=np.float32)
    out = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
    experimental.direct_BP3D(pid, out, torch.from_numpy(y).cuda())
    refb = bp(y, ov, (nz, ny, nx), vs)
    if np.linalg.norm(refb) > 0:
        assert rel_l2(out.cpu().numpy(), refb) <= 1e-5
    # the matched adjoint on the same shape: FP64 dot test
    from astra import autograd as AG
    xt = torch.from_numpy(vol).cuda()
    yt = torch.from_numpy(y).cuda()
    lhs = float((AG.forward_project(xt, pid).double() * yt.double()).sum())
    rhs = float((xt.double() * AG.back_project(yt, pid).double()).sum())
    if abs(lhs) > 0:
        assert abs(lhs - rhs) / abs(lhs) <= 1e-5


SHAPES_2D = [(1, 9, 7, 11), (9, 1, 7, 11), (8, 8, 1, 11), (8, 8, 13, 1), (1, 1, 3, 3), (17, 5, 23, 13)]


@pytest.mark.parametrize("kind", ["parallel", "fanflat"])
@pytest.mark.parametrize("shape", SHAPES_2D)
def test_2d_edge_shapes(kind, shape):
    import astra
    from oracle import geometry as OG, proj2d as O2
    w, h, na, nd = shape
    vg = astra.create_vol_geom(h, w)
    ang = np.linspace(0.2, 0.2 + np.pi, na, endpoint=False)
    if kind == "parallel":
        pg = astra.create_proj_geom("parallel", 1.0, nd, ang)
        ov, osc = OG.normalise_2d(OG.gen_par_projections(ang, nd, 1.0), "par", vg)
        fp, bp = O2.par_fp, O2.par_bp
    else:
        pg = astra.create_proj_geom("fanflat", 1.0, nd, ang, 30.0, 15.0)
        ov, osc = OG.normalise_2d(OG.gen_fan_projections(ang, nd, 30.0, 15.0, 1.0), "fan", vg)
        fp, bp = O2.fan_fp, O2.fan_bp
    pid = astra.create_projector("cuda", pg, vg)
    rng = np.random.default_rng(sum(shape) + 1)
    vol = rng.random((h, w), dtype=np.float32)
    _, got = astra.create_sino(vol, pid)
    ref = fp(vol, ov, nd, osc)
    if np.linalg.norm(ref) > 0:
        assert rel_l2(got, ref) <= 1e-5
    y = rng.random((na, nd), dtype=np.float32)
    _, gotb = astra.create_backprojection(y, pid)
    refb = bp(y, ov, (h, w), osc)
    if np.linalg.norm(refb) > 0:
        assert rel_l2(gotb, refb) <= 1e-5


def test_filtered_algorithms_with_one_detector():
    import astra
    from oracle import fdk as OF, geometry as OG
    # FBP with one detector bin: N = 2
    vg = astra.create_vol_geom(4, 4)
    ang = np.linspace(0, np.pi, 8, endpoint=False)
    pg = astra.create_proj_geom("parallel", 1.0, 1, ang)
    ov, osc = OG.normalise_2d(OG.gen_par_projections(ang, 1, 1.0), "par", vg)
    sino = np.random.default_rng(3).random((8, 1), dtype=np.float32)
    pid = astra.create_projector("cuda", pg, vg)
    sid = astra.data2d.create("-sino", pg, sino)
    rid = astra.data2d.create("-vol", vg, 0)
    cfg = astra.astra_dict("FBP_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
    astra.algorithm.run(astra.algorithm.create(cfg))
    ref = OF.fbp_par(sino, ov, (4, 4), osc, 1.0)
    assert rel_l2(astra.data2d.get(rid), ref) <= 1e-5
