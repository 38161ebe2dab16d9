"""The 3D voxel-driven BP launched in angle chunks.

Above 8 GiB of padded projections the BP runs its angles in chunks of 40
launches (one volume read-modify-write per chunk; update modes sum the
earlier chunks in a scratch volume that the last chunk adds before its
epilogue).  ASTRA_AMD_BP3D_ACHUNK forces the chunk size at small sizes, so
every chunked form is checked here: SET and ADD, the cone default / FDK /
2D-weighted kernels, par3d (per-angle weights), and SIRT3D's fused update
with a box constraint.  Each result is held to the FP64 oracle at the
single-BP bar (rel-L2 <= 1e-5; <= 1e-4 after SIRT iterations) and to the
unchunked GPU result, which differs only in the fp32 order of the sums.
"""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

ENV = "ASTRA_AMD_BP3D_ACHUNK"


@pytest.fixture
def chunked():
    """Sets the chunk size for the calls inside the test (7 angles: five
    full chunks and a partial one over 37 angles)."""
    old = os.environ.get(ENV)

    def set_chunk(n):
        if n:
            os.environ[ENV] = str(n)
        else:
            os.environ.pop(ENV, None)
    yield set_chunk
    set_chunk(old)


def _case(kind):
    import astra
    from oracle import geometry as OG
    na, nu, nv = 37, 44, 40
    vg = astra.create_vol_geom(36, 40, 28, -20, 20, -18, 18, -14, 14)
    if kind == "cone":
        ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("cone", 1.0, 1.0, nv, nu, ang, 150.0, 80.0)
        ovec, vs = OG.normalise_3d(OG.gen_cone_projections(ang, nu, nv, 150.0, 80.0, 1.0, 1.0),
                                   "cone", vg)
    else:
        ang = np.linspace(0, np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("parallel3d", 1.0, 1.0, nv, nu, ang)
        ovec, vs = OG.normalise_3d(OG.gen_par3d_projections(ang, nu, nv, 1.0, 1.0), "par3d", vg)
    return astra, vg, pg, ovec, vs, (nv, na, nu), (28, 36, 40)


def _bp(pid, proj, init, mode, kernel):
    import torch
    import astra
    from astra import _core
    p = astra.projector3d.get(pid)
    pt = torch.from_numpy(proj).cuda()
    vt = torch.from_numpy(init.copy()).cuda()
    _core.bp(p, pt, vt, mode, kernel=kernel)
    return vt.cpu().numpy()


@pytest.mark.parametrize("kind,kernel", [("cone", "default"), ("cone", "fdk"), ("cone", "2d"),
                                         ("par3d", "default")])
@pytest.mark.parametrize("mode", ["set", "add"])
def test_chunked_bp(chunked, kind, kernel, mode):
    from astra import _lib
    from oracle import proj3d as O3
    astra, vg, pg, ovec, vs, pshape, vshape = _case(kind)
    pid = astra.create_projector("cuda3d", pg, vg)
    rng = np.random.default_rng(12)
    proj = rng.random(pshape, dtype=np.float32)
    init = rng.random(vshape, dtype=np.float32) if mode == "add" else np.full(vshape, 5.0, np.float32)
    km = {"default": _lib.KER3D_DEFAULT, "fdk": _lib.KER3D_FDK_WEIGHTING, "2d": _lib.KER3D_2D_WEIGHTING}
    m = _lib.MODE_ADD if mode == "add" else _lib.MODE_SET
    chunked(0)
    whole = _bp(pid, proj, init, m, km[kernel])
    chunked(7)
    parts = _bp(pid, proj, init, m, km[kernel])
    chunked(0)
    if kind == "cone":
        ref = O3.cone_bp(proj, ovec, vshape, vs, kernel=kernel)
    else:
        ref = O3.par3d_bp(proj, ovec, vshape, vs)
    if mode == "add":
        ref = ref + init
    e_parts, e_whole = rel_l2(parts, ref), rel_l2(whole, ref)
    print(kind, kernel, mode, "chunked", e_parts, "whole", e_whole)
    assert e_parts <= 1e-5 and e_whole <= 1e-5
    assert rel_l2(parts, whole) <= 1e-6
    astra.projector3d.delete(pid)


@pytest.mark.parametrize("kind", ["cone", "par3d"])
def test_chunked_sirt_update(chunked, kind):
    """SIRT3D's BP epilogue (x = clamp(x + w * BP(r))) over chunked angles."""
    import torch
    from astra import phantom
    from oracle import algorithms as OA, proj3d as O3
    astra, vg, pg, ovec, vs, pshape, vshape = _case(kind)
    pid = astra.create_projector("cuda3d", pg, vg)
    nv, na, nu = pshape
    if kind == "cone":
        fp = lambda v: O3.cone_fp(v, ovec, nu, nv, vs)
        bp = lambda s: O3.cone_bp(s, ovec, vshape, vs)
    else:
        fp = lambda v: O3.par3d_fp(v, ovec, nu, nv, vs)
        bp = lambda s: O3.par3d_bp(s, ovec, vshape, vs)
    truth = phantom.ellipsoids_3d(40, seed=13)[:28, 2:38, :40].copy()
    sino = fp(truth).astype(np.float32)
    iters = 4
    outs = {}
    for c in (0, 7):
        chunked(c)
        st = torch.from_numpy(sino).cuda()
        rt = torch.zeros(vshape, dtype=torch.float32, device="cuda")
        sid = astra.data3d.link("-sino", pg, st)
        rid = astra.data3d.link("-vol", vg, rt)
        cfg = astra.astra_dict("SIRT3D_CUDA")
        cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid,
                   option={"MinConstraint": 0.0, "MaxConstraint": 0.6})
        aid = astra.algorithm.create(cfg)
        astra.algorithm.run(aid, iters)
        outs[c] = rt.cpu().numpy()
        astra.algorithm.delete(aid)
        astra.data3d.delete([sid, rid])
    chunked(0)
    ref = OA.sirt(fp, bp, sino, np.zeros(vshape), iters, min_c=0.0, max_c=0.6)
    e7, e0 = rel_l2(outs[7], ref), rel_l2(outs[0], ref)
    print(kind, "SIRT chunked", e7, "whole", e0)
    assert e7 <= 1e-4 and e0 <= 1e-4
    assert rel_l2(outs[7], outs[0]) <= 1e-5
    assert outs[7].min() >= 0.0 and outs[7].max() <= 0.6
    astra.projector3d.delete(pid)
