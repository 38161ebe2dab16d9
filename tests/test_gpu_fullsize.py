"""Parity at the full workload sizes.

2D fan: 1024^2 volume, 720 angles over [0, 2pi), 1536 detectors, source-
origin 1000, origin-detector 500, Poisson noise at 1e5 photons, FBP and
200 SIRT iterations.  3D cone: 1024^3 volume, 1440 angles, 1536^2 detector
(SOD 1000, ODD 500), unchunked.

The FP64 oracle cannot sweep these sizes in seconds, so parity is checked
at full spatial size on a subset of angles (the kernels treat every angle
independently) or on a detector / volume window (a cone-vector geometry of
the window's rays, a windowed volume geometry), and the full-angle runs are
checked through properties that do not depend on size: linearity,
run composition, residual decrease and the reference's scaling test
(reconstructed density of a uniform object = 1 within 5e-2,
tests/python/functional/test_rec_scaling.py:89-145).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

N2, NA2, ND2, SOD, ODD = 1024, 720, 1536, 1000.0, 500.0
SUBSET = [0, 97, 181, 360, 455, 539, 600, 719]


def _algo(astra, atype, pid, sid, rid, iters, **opts):
    cfg = astra.astra_dict(atype)
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
    if opts:
        cfg["option"] = opts
    aid = astra.algorithm.create(cfg)
    astra.algorithm.run(aid, iters)
    return aid


@pytest.fixture(scope="module")
def fan1024():
    import astra
    from astra import phantom
    angles = np.linspace(0, 2 * np.pi, NA2, endpoint=False)
    vg = astra.create_vol_geom(N2, N2)
    pg = astra.create_proj_geom("fanflat", 1.0, ND2, angles, SOD, ODD)
    pid = astra.create_projector("cuda", pg, vg)
    vol = phantom.ellipses_2d(N2, seed=2)
    _, sino = astra.create_sino(vol, pid)
    noisy = astra.add_noise_to_sino(sino, 1e5, seed=2)
    return dict(astra=astra, angles=angles, vg=vg, pg=pg, pid=pid, vol=vol, sino=sino, noisy=noisy)


def _subset_geoms(c):
    from oracle import geometry as OG
    ang = c["angles"][SUBSET]
    pg = c["astra"].create_proj_geom("fanflat", 1.0, ND2, ang, SOD, ODD)
    ovec, osc = OG.normalise_2d(OG.gen_fan_projections(ang, ND2, SOD, ODD, 1.0), "fan", c["vg"])
    return pg, ovec, osc


def test_fan1024_fp_bp_vs_oracle_on_angle_subset(fan1024):
    from oracle import proj2d as O2
    c = fan1024
    astra = c["astra"]
    pg, ovec, osc = _subset_geoms(c)
    pid = astra.create_projector("cuda", pg, c["vg"])
    _, got = astra.create_sino(c["vol"], pid)
    ref = O2.fan_fp(c["vol"], ovec, ND2, osc)
    assert rel_l2(got, ref) <= 1e-5
    # the full-angle sinogram holds the same rows
    assert rel_l2(c["sino"][SUBSET], ref) <= 1e-5
    # BP of sinogram data: 1e-5.  White noise (every texel independent) is
    # the worst case for coordinate rounding: detector coordinates up to
    # 1540 carry fp32 steps of 6e-5 texels (the reference's fp32 formula
    # has the same), which white noise turns into ~3e-5 relative error.
    y = np.ascontiguousarray(c["sino"][SUBSET])
    _, gbp = astra.create_backprojection(y, pid)
    assert rel_l2(gbp, O2.fan_bp(y, ovec, (N2, N2), osc)) <= 1e-5
    w = np.random.default_rng(3).random((len(SUBSET), ND2), dtype=np.float32)
    _, gbw = astra.create_backprojection(w, pid)
    assert rel_l2(gbw, O2.fan_bp(w, ovec, (N2, N2), osc)) <= 1e-4


def test_fan1024_fbp(fan1024):
    from oracle import fdk as OF
    c = fan1024
    astra = c["astra"]
    # parity on the angle subset (filter + FBP-weighted BP, same code path)
    pg, ovec, osc = _subset_geoms(c)
    pid = astra.create_projector("cuda", pg, c["vg"])
    sub = np.ascontiguousarray(c["noisy"][SUBSET])
    sid = astra.data2d.create("-sino", pg, sub)
    rid = astra.data2d.create("-vol", c["vg"], 0)
    _algo(astra, "FBP_CUDA", pid, sid, rid, 1)
    ref = OF.fbp_fan(sub, ovec, (N2, N2), osc, 1.0)
    # The ramp-filtered noisy sinogram is dominated by texel-to-texel
    # jumps, so the BP's fp32 detector coordinate (values up to 1540, steps
    # of 6e-5 texels; the reference evaluates the same fp32 formula) sets
    # the error: rounding only the final coordinate of the FP64 oracle to
    # fp32 already gives 4.7e-5 here; the fp32 evaluation gives ~2e-4.
    assert rel_l2(astra.data2d.get(rid), ref) <= 5e-4
    # full 720 angles: the scaling test on a uniform disk, noise-free and noisy
    yy, xx = np.mgrid[0:N2, 0:N2] - (N2 - 1) / 2.0
    disk = ((xx ** 2 + yy ** 2) <= (0.4 * N2) ** 2).astype(np.float32)
    inner = (xx ** 2 + yy ** 2) <= (0.3 * N2) ** 2
    _, dsino = astra.create_sino(disk, c["pid"])
    for s in (dsino, astra.add_noise_to_sino(dsino, 1e5, seed=5)):
        sid = astra.data2d.create("-sino", c["pg"], s)
        rid = astra.data2d.create("-vol", c["vg"], 0)
        _algo(astra, "FBP_CUDA", c["pid"], sid, rid, 1)
        rec = astra.data2d.get(rid)
        assert abs(rec[inner].mean() - 1.0) <= 5e-2, rec[inner].mean()


def test_fan1024_sirt200(fan1024):
    """200 SIRT iterations on the noisy 1024^2 sinogram: run(100); run(100)
    equals run(200) bit for bit, the residual norm decreases, and the
    reconstruction approaches the phantom."""
    c = fan1024
    astra = c["astra"]
    sid = astra.data2d.create("-sino", c["pg"], c["noisy"])
    r1 = astra.data2d.create("-vol", c["vg"], 0)
    a1 = _algo(astra, "SIRT_CUDA", c["pid"], sid, r1, 100)
    res100 = astra.algorithm.get_res_norm(a1)
    err100 = rel_l2(astra.data2d.get(r1), c["vol"])
    astra.algorithm.run(a1, 100)
    res200 = astra.algorithm.get_res_norm(a1)
    x1 = astra.data2d.get(r1)
    r2 = astra.data2d.create("-vol", c["vg"], 0)
    _algo(astra, "SIRT_CUDA", c["pid"], sid, r2, 200)
    assert np.array_equal(x1, astra.data2d.get(r2))
    assert res200 < res100
    err200 = rel_l2(x1, c["vol"])
    assert err200 < err100 < 1.0


# ---------------------------------------------------------------- 3D large
N3, NA3, ND3 = 1024, 1440, 1536
ANG3 = np.linspace(0, 2 * np.pi, NA3, endpoint=False)
PICK = [0, 100, 333]  # away from the 45-degree direction switch


@pytest.fixture(scope="module")
def cone1024():
    import torch
    import astra
    from astra import phantom
    vg = astra.create_vol_geom(N3, N3, N3)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, ND3, ND3, ANG3, SOD, ODD)
    pid = astra.create_projector("cuda3d", pg, vg)
    vol = phantom.ellipsoids_3d_device(N3, count=20, seed=0, device="cuda")
    yield dict(astra=astra, vg=vg, pg=pg, pid=pid, vol=vol)
    astra.projector3d.delete(pid)
    del vol
    torch.cuda.empty_cache()


def test_cone1024_fp_window_vs_oracle(cone1024):
    """Full 1440-angle FP of the 1024^3 phantom; the rays of a 48 x 40
    detector window at three angles against the oracle (a cone-vector
    geometry of exactly those rays)."""
    import torch
    from astra import experimental
    from oracle import geometry as OG, proj3d as O3
    c = cone1024
    proj = torch.empty((ND3, NA3, ND3), dtype=torch.float32, device="cuda")
    experimental.direct_FP3D(c["pid"], c["vol"], proj)
    u0, v0, wu, wv = 700, 730, 48, 40
    full, _ = OG.normalise_3d(OG.gen_cone_projections(ANG3[PICK], ND3, ND3, SOD, ODD, 1.0, 1.0), "cone",
                              OG.vol_geom_3d(N3, N3, N3))
    win = full.copy()
    win[:, 3:6] = full[:, 3:6] + u0 * full[:, 6:9] + v0 * full[:, 9:12]
    vol = c["vol"].cpu().numpy()
    ref = O3.cone_fp(vol, win, wu, wv)
    got = proj[v0:v0 + wv][:, PICK][:, :, u0:u0 + wu].cpu().numpy()
    assert rel_l2(got, ref) <= 1e-5
    # linearity of the full FP: A(2 x) = 2 A x exactly (power-of-two scale)
    proj2 = torch.empty_like(proj)
    experimental.direct_FP3D(c["pid"], c["vol"] * 2, proj2)
    assert torch.equal(proj2, proj * 2)
    del proj, proj2
    torch.cuda.empty_cache()


def test_cone1024_bp_block_vs_oracle(cone1024):
    """BP over the full 1024^3 volume from 1536^2 projections (three
    angles); a 12^3 voxel block against the oracle on a windowed volume
    geometry of exactly those voxels."""
    import torch
    import astra
    from astra import experimental
    from oracle import geometry as OG, proj3d as O3
    pg = astra.create_proj_geom("cone", 1.0, 1.0, ND3, ND3, ANG3[PICK], SOD, ODD)
    pid = astra.create_projector("cuda3d", pg, cone1024["vg"])
    # smooth projection data (detector coordinates up to 1540 carry fp32
    # steps of 6e-5 texels, so white noise would measure rounding, not
    # sampling): the phantom's projections at the three angles
    y = torch.empty((ND3, len(PICK), ND3), dtype=torch.float32, device="cuda")
    experimental.direct_FP3D(pid, cone1024["vol"], y)
    out = torch.empty((N3, N3, N3), dtype=torch.float32, device="cuda")
    experimental.direct_BP3D(pid, out, y)
    x0, y0, z0, b = 300, 610, 505, 12
    vg = OG.vol_geom_3d(b, b, b)
    vg["option"] = {"WindowMinX": x0 - N3 / 2, "WindowMaxX": x0 + b - N3 / 2,
                    "WindowMinY": y0 - N3 / 2, "WindowMaxY": y0 + b - N3 / 2,
                    "WindowMinZ": z0 - N3 / 2, "WindowMaxZ": z0 + b - N3 / 2}
    v, vs = OG.normalise_3d(OG.gen_cone_projections(ANG3[PICK], ND3, ND3, SOD, ODD, 1.0, 1.0), "cone", vg)
    ref = O3.cone_bp(y.cpu().numpy(), v, (b, b, b), vs)
    got = out[z0:z0 + b, y0:y0 + b, x0:x0 + b].cpu().numpy()
    assert rel_l2(got, ref) <= 1e-5
    astra.projector3d.delete(pid)
    del out
    torch.cuda.empty_cache()


def test_cone1024_fdk_scaling(cone1024):
    """FDK of a uniform ball at 1024^3 / 1440 x 1536^2, unchunked: the
    central slab reconstructs density 1 within 5e-2."""
    import torch
    from astra import _lib, experimental, projector3d
    from astra.algorithm import fdk
    c = cone1024
    ar = torch.arange(N3, device="cuda", dtype=torch.float32) - (N3 - 1) / 2.0
    r2 = ar.view(-1, 1, 1) ** 2 + ar.view(1, -1, 1) ** 2 + ar.view(1, 1, -1) ** 2
    ball = (r2 <= (0.35 * N3) ** 2).to(torch.float32)
    proj = torch.empty((ND3, NA3, ND3), dtype=torch.float32, device="cuda")
    experimental.direct_FP3D(c["pid"], ball, proj)
    rec = torch.empty_like(ball)
    fdk(projector3d.get(c["pid"]), proj, rec, _lib.MODE_SET)
    zc = slice(N3 // 2 - 32, N3 // 2 + 32)
    inner = r2[zc] <= (0.25 * N3) ** 2
    m = float(rec[zc][inner].mean())
    assert abs(m - 1.0) <= 5e-2, m
    del proj, rec, ball, r2
    torch.cuda.empty_cache()
