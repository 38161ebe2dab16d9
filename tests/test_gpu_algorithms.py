"""GPU parity of the reconstruction algorithms against the FP64 oracle.

Bar: relative L2 <= 1e-4 after N SIRT / CGLS iterations, <= 1e-5 for the
one-shot FDK / FBP (filter + one BP).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _run_algo(astra, atype, pid, sino_id, rec_id, iters, **opts):
    cfg = astra.astra_dict(atype)
    cfg["ProjectorId"] = pid
    cfg["ProjectionDataId"] = sino_id
    cfg["ReconstructionDataId"] = rec_id
    if opts:
        cfg["option"] = opts
    aid = astra.algorithm.create(cfg)
    astra.algorithm.run(aid, iters)
    return aid


@pytest.fixture(scope="module")
def fan2d():
    import astra
    from astra import phantom
    from oracle import geometry as OG, proj2d as O2
    n, na, nd = 64, 90, 96
    angles = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n)
    pg = astra.create_proj_geom("fanflat", 1.0, nd, angles, 200.0, 100.0)
    ovec, osc = OG.normalise_2d(OG.gen_fan_projections(angles, nd, 200.0, 100.0, 1.0), "fan", vg)
    vol = phantom.ellipses_2d(n, seed=21)
    sino = O2.fan_fp(vol, ovec, nd, osc).astype(np.float32)
    return dict(astra=astra, vg=vg, pg=pg, n=n, na=na, nd=nd, sino=sino,
                fp=lambda v: O2.fan_fp(v, ovec, nd, osc),
                bp=lambda s: O2.fan_bp(s, ovec, (n, n), osc), ovec=ovec, osc=osc)


@pytest.mark.parametrize("algo", ["SIRT_CUDA", "CGLS_CUDA"])
def test_2d_iterative(fan2d, algo):
    from oracle import algorithms as OA
    c = fan2d
    astra = c["astra"]
    pid = astra.create_projector("cuda", c["pg"], c["vg"])
    sid = astra.data2d.create("-sino", c["pg"], c["sino"])
    rid = astra.data2d.create("-vol", c["vg"], 0)
    iters = 20
    aid = _run_algo(astra, algo, pid, sid, rid, iters)
    got = astra.data2d.get(rid)
    x0 = np.zeros((c["n"], c["n"]))
    if algo == "SIRT_CUDA":
        ref = OA.sirt(c["fp"], c["bp"], c["sino"], x0, iters)
    else:
        ref = OA.cgls(c["fp"], c["bp"], c["sino"], x0, iters)
    e = rel_l2(got, ref)
    print(algo, "rel-L2 after", iters, e)
    assert e <= 1e-4
    # residual norm matches ||b - A x|| of the oracle at the GPU iterate
    rn = astra.algorithm.get_res_norm(aid)
    assert abs(rn - OA.residual_norm(c["fp"], c["sino"], got)) <= 1e-4 * np.linalg.norm(c["sino"])


def test_sirt_constraints_and_runs_compose(fan2d):
    """MinConstraint / MaxConstraint bound the iterate exactly and
    run(a); run(b) == run(a + b) for SIRT (src/CudaSirtAlgorithm.cpp:328-336)."""
    c = fan2d
    astra = c["astra"]
    pid = astra.create_projector("cuda", c["pg"], c["vg"])
    sid = astra.data2d.create("-sino", c["pg"], c["sino"])
    r1 = astra.data2d.create("-vol", c["vg"], 0)
    a1 = _run_algo(astra, "SIRT_CUDA", pid, sid, r1, 0, MinConstraint=0.1, MaxConstraint=0.6)
    astra.algorithm.run(a1, 3)
    astra.algorithm.run(a1, 4)
    r2 = astra.data2d.create("-vol", c["vg"], 0)
    _run_algo(astra, "SIRT_CUDA", pid, sid, r2, 7, MinConstraint=0.1, MaxConstraint=0.6)
    x1, x2 = astra.data2d.get(r1), astra.data2d.get(r2)
    assert x1.min() >= 0.1 and x1.max() <= 0.6
    assert np.array_equal(x1, x2)


def test_fbp_fan_and_par(fan2d):
    from oracle import fdk as OF, geometry as OG
    c = fan2d
    astra = c["astra"]
    pid = astra.create_projector("cuda", c["pg"], c["vg"])
    sid = astra.data2d.create("-sino", c["pg"], c["sino"])
    rid = astra.data2d.create("-vol", c["vg"], 0)
    _run_algo(astra, "FBP_CUDA", pid, sid, rid, 1)
    got = astra.data2d.get(rid)
    ref = OF.fbp_fan(c["sino"], c["ovec"], (c["n"], c["n"]), c["osc"], 1.0)
    e = rel_l2(got, ref)
    print("FBP fan rel-L2", e)
    assert e <= 1e-5
    # parallel
    n, na, nd = 64, 90, 96
    angles = np.linspace(0, np.pi, na, endpoint=False)
    pg = astra.create_proj_geom("parallel", 1.0, nd, angles)
    pid2 = astra.create_projector("cuda", pg, c["vg"])
    ovec, osc = OG.normalise_2d(OG.gen_par_projections(angles, nd, 1.0), "par", c["vg"])
    sino = np.random.default_rng(22).random((na, nd), dtype=np.float32)
    sid2 = astra.data2d.create("-sino", pg, sino)
    rid2 = astra.data2d.create("-vol", c["vg"], 0)
    _run_algo(astra, "FBP_CUDA", pid2, sid2, rid2, 1)
    ref2 = OF.fbp_par(sino, ovec, (n, n), osc, 1.0)
    assert rel_l2(astra.data2d.get(rid2), ref2) <= 1e-5


@pytest.fixture(scope="module")
def cone3d():
    import astra
    from astra import phantom
    from oracle import geometry as OG, proj3d as O3
    n, na, nd = 32, 45, 48
    angles = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, angles, 150.0, 75.0)
    ovec, vs = OG.normalise_3d(OG.gen_cone_projections(angles, nd, nd, 150.0, 75.0, 1.0, 1.0),
                               "cone", OG.vol_geom_3d(n, n, n))
    vol = phantom.ellipsoids_3d(n, seed=23)
    proj = O3.cone_fp(vol, ovec, nd, nd, vs).astype(np.float32)
    return dict(astra=astra, vg=vg, pg=pg, n=n, na=na, nd=nd, proj=proj, ovec=ovec, vs=vs,
                fp=lambda v: O3.cone_fp(v, ovec, nd, nd, vs),
                bp=lambda p: O3.cone_bp(p, ovec, (n, n, n), vs))


@pytest.mark.parametrize("algo", ["SIRT3D_CUDA", "CGLS3D_CUDA"])
def test_3d_iterative(cone3d, algo):
    from oracle import algorithms as OA
    c = cone3d
    astra = c["astra"]
    sid = astra.data3d.create("-sino", c["pg"], c["proj"])
    rid = astra.data3d.create("-vol", c["vg"], 0)
    cfg = astra.astra_dict(algo)
    cfg["ProjectionDataId"] = sid
    cfg["ReconstructionDataId"] = rid
    aid = astra.algorithm.create(cfg)
    iters = 8
    astra.algorithm.run(aid, iters)
    got = astra.data3d.get(rid)
    x0 = np.zeros((c["n"],) * 3)
    if algo == "SIRT3D_CUDA":
        ref = OA.sirt(c["fp"], c["bp"], c["proj"], x0, iters)
    else:
        ref = OA.cgls(c["fp"], c["bp"], c["proj"], x0, iters)
    e = rel_l2(got, ref)
    print(algo, "rel-L2 after", iters, e)
    assert e <= 1e-4


def test_fdk(cone3d):
    from oracle import fdk as OF
    c = cone3d
    astra = c["astra"]
    sid = astra.data3d.create("-sino", c["pg"], c["proj"])
    rid = astra.data3d.create("-vol", c["vg"], 0)
    cfg = astra.astra_dict("FDK_CUDA")
    cfg["ProjectionDataId"] = sid
    cfg["ReconstructionDataId"] = rid
    aid = astra.algorithm.create(cfg)
    astra.algorithm.run(aid)
    got = astra.data3d.get(rid)
    ref = OF.fdk(c["proj"], c["ovec"], (c["n"],) * 3, c["vs"])
    e = rel_l2(got, ref)
    print("FDK rel-L2", e)
    assert e <= 1e-5
    # the input projections are not modified
    assert np.array_equal(astra.data3d.get(sid), c["proj"])


@pytest.mark.parametrize("kind", ["parallel", "fanflat"])
def test_em(kind):
    """EM_CUDA against the oracle's EM over the oracle projectors."""
    import astra
    from astra import phantom
    from oracle import algorithms as OA, geometry as OG, proj2d as O2
    n, na, nd = 48, 60, 72
    vg = astra.create_vol_geom(n, n)
    if kind == "parallel":
        ang = np.linspace(0, np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("parallel", 1.0, nd, ang)
        v, osc = OG.normalise_2d(OG.gen_par_projections(ang, nd, 1.0), "parallel", vg)
        fp, bp = (lambda x: O2.par_fp(x, v, nd, osc)), (lambda y: O2.par_bp(y, v, (n, n), osc))
    else:
        ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("fanflat", 1.0, nd, ang, 150.0, 75.0)
        v, osc = OG.normalise_2d(OG.gen_fan_projections(ang, nd, 150.0, 75.0, 1.0), "fan", vg)
        fp, bp = (lambda x: O2.fan_fp(x, v, nd, osc)), (lambda y: O2.fan_bp(y, v, (n, n), osc))
    sino = fp(phantom.ellipses_2d(n, seed=63)).astype(np.float32)
    pid = astra.create_projector("cuda", pg, vg)
    sid = astra.data2d.create("-sino", pg, sino)
    rid = astra.data2d.create("-vol", vg, 1.0)
    aid = _run_algo(astra, "EM_CUDA", pid, sid, rid, 10)
    ref = OA.em(fp, bp, sino, np.ones((n, n)), 10)
    assert rel_l2(astra.data2d.get(rid), ref) <= 1e-4
    rn = astra.algorithm.get_res_norm(aid)
    assert abs(rn - OA.residual_norm(fp, sino, astra.data2d.get(rid))) <= 1e-4 * np.linalg.norm(sino)


@pytest.mark.parametrize("kind", ["parallel", "fanflat"])
def test_sart(kind):
    """SART_CUDA (custom order, constraint, two runs continuing the order)
    against the oracle's SART with SART's unweighted BP."""
    import astra
    from astra import phantom
    from oracle import algorithms as OA, geometry as OG, proj2d as O2
    n, na, nd = 40, 30, 60
    vg = astra.create_vol_geom(n, n)
    if kind == "parallel":
        ang = np.linspace(0, np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("parallel", 1.0, nd, ang)
        v, osc = OG.normalise_2d(OG.gen_par_projections(ang, nd, 1.0), "parallel", vg)
        fp1 = lambda x, a: O2.par_fp(x, v[a:a + 1], nd, osc)[0]
        full = lambda x: O2.par_fp(x, v, nd, osc)
        okind = "parallel"
    else:
        ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("fanflat", 1.0, nd, ang, 120.0, 60.0)
        v, osc = OG.normalise_2d(OG.gen_fan_projections(ang, nd, 120.0, 60.0, 1.0), "fan", vg)
        fp1 = lambda x, a: O2.fan_fp(x, v[a:a + 1], nd, osc)[0]
        full = lambda x: O2.fan_fp(x, v, nd, osc)
        okind = "fan"
    bp1 = lambda r, a: O2.sart_bp(r, v[a], (n, n), okind, osc)
    sino = full(phantom.ellipses_2d(n, seed=64)).astype(np.float32)
    order = [int(i) for i in np.random.default_rng(65).permutation(na)]
    pid = astra.create_projector("cuda", pg, vg)
    sid = astra.data2d.create("-sino", pg, sino)
    rid = astra.data2d.create("-vol", vg, 0)
    aid = _run_algo(astra, "SART_CUDA", pid, sid, rid, 25, ProjectionOrder="custom",
                    ProjectionOrderList=order, Relaxation=0.8, MinConstraint=0.0)
    astra.algorithm.run(aid, 20)  # continues the order at iteration 25
    seq = [order[k % na] for k in range(45)]
    ref = OA.sart(fp1, bp1, sino, np.zeros((n, n)), seq, relaxation=0.8, min_c=0.0,
                  line_w=OA._inv(full(np.ones((n, n)))))
    assert rel_l2(astra.data2d.get(rid), ref) <= 1e-4
