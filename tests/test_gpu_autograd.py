"""Differentiable projectors on the GPU: matched adjoints against the FP64
oracle transposes, dot-product identity, finite differences, the
reference's unmatched BP as a reference gradient, and unrolled SIRT with
recomputation against autograd through the stored iterates."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _cone(n, na, nd, sod=None, odd=None):
    import astra
    sod = 4.0 * n if sod is None else sod
    odd = 2.0 * n if odd is None else odd
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, ang, sod, odd)
    return astra.create_projector("cuda3d", pg, vg), pg, vg


def _rand(shape, seed):
    import torch
    g = torch.Generator().manual_seed(seed)
    return torch.rand(shape, generator=g, dtype=torch.float32).cuda()


def test_matched_bp_equals_oracle_transpose_3d():
    import astra
    from astra import autograd as AG
    from oracle import geometry as OG, proj3d as O3
    n, na, nd = 24, 30, 32
    pid, pg, vg = _cone(n, na, nd)
    y = _rand((nd, na, nd), 1)
    got = AG.back_project(y, pid).cpu().numpy()
    v, vs = OG.normalise_3d(OG.gen_cone_projections(pg["ProjectionAngles"], nd, nd, 4.0 * n, 2.0 * n, 1.0, 1.0),
                            "cone", OG.vol_geom_3d(n, n, n))
    ref = O3.cone_fp_T(y.cpu().numpy().astype(np.float64), v, (n, n, n), vs)
    assert rel_l2(got, ref) <= 1e-5
    # par3d
    pg3 = astra.create_proj_geom("parallel3d", 1.0, 1.0, nd, nd, pg["ProjectionAngles"][:12])
    pid3 = astra.create_projector("cuda3d", pg3, vg)
    y3 = _rand((nd, 12, nd), 2)
    got3 = AG.back_project(y3, pid3).cpu().numpy()
    v3, vs3 = OG.normalise_3d(OG.gen_par3d_projections(pg["ProjectionAngles"][:12], nd, nd, 1.0, 1.0),
                              "parallel3d", OG.vol_geom_3d(n, n, n))
    ref3 = O3.par3d_fp_T(y3.cpu().numpy().astype(np.float64), v3, (n, n, n), vs3)
    assert rel_l2(got3, ref3) <= 1e-5


@pytest.mark.parametrize("kind", ["parallel", "fanflat"])
def test_matched_bp_equals_oracle_transpose_2d(kind):
    import astra
    from astra import autograd as AG
    from oracle import geometry as OG, proj2d as O2
    n, na, nd = 40, 45, 60
    vg = astra.create_vol_geom(n, n)
    if kind == "parallel":
        ang = np.linspace(0, np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("parallel", 1.0, nd, ang)
        v, osc = OG.normalise_2d(OG.gen_par_projections(ang, nd, 1.0), "parallel", OG.vol_geom_2d(n, n))
        T = O2.par_fp_T
    else:
        ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
        pg = astra.create_proj_geom("fanflat", 1.0, nd, ang, 3.0 * n, 1.5 * n)
        v, osc = OG.normalise_2d(OG.gen_fan_projections(ang, nd, 3.0 * n, 1.5 * n, 1.0), "fan",
                                 OG.vol_geom_2d(n, n))
        T = O2.fan_fp_T
    pid = astra.create_projector("cuda", pg, vg)
    y = _rand((na, nd), 3)
    got = AG.back_project(y, pid).cpu().numpy()
    ref = T(y.cpu().numpy().astype(np.float64), v, (n, n), osc)
    assert rel_l2(got, ref) <= 1e-5


def test_dot_product_identity_64cube():
    """<A x, y> = <x, A^T y> at 64^3 / 90 angles / 96^2, sums in FP64 (the
    fp32 kernels leave ~1e-7 relative rounding)."""
    from astra import autograd as AG
    pid, _, _ = _cone(64, 90, 96)
    x = _rand((64, 64, 64), 4)
    y = _rand((96, 90, 96), 5)
    ax = AG.forward_project(x, pid).double()
    aty = AG.back_project(y, pid).double()
    lhs = float((ax * y.double()).sum())
    rhs = float((x.double() * aty).sum())
    assert abs(lhs - rhs) / abs(lhs) <= 1e-5, (lhs, rhs)


def test_gradient_finite_differences():
    """d/dx of L(x) = 1/2 ||A x - y||^2 (a quadratic: central differences are
    exact up to rounding) along random directions, loss summed in FP64."""
    import torch
    from astra import autograd as AG
    pid, _, _ = _cone(32, 40, 48)
    y = _rand((48, 40, 48), 6) * 20
    x = _rand((32, 32, 32), 7).requires_grad_(True)
    loss = 0.5 * ((AG.forward_project(x, pid).double() - y.double()) ** 2).sum()
    (g,) = torch.autograd.grad(loss, x)
    for seed in (8, 9, 10):
        d = _rand((32, 32, 32), seed) - 0.5
        eps = 0.5
        with torch.no_grad():
            lp = 0.5 * ((AG.forward_project(x + eps * d, pid).double() - y.double()) ** 2).sum()
            lm = 0.5 * ((AG.forward_project(x - eps * d, pid).double() - y.double()) ** 2).sum()
        fd = float(lp - lm) / (2 * eps)
        an = float((g.double() * d.double()).sum())
        assert abs(fd - an) <= 1e-3 * abs(an), (seed, fd, an)


def test_gradient_close_to_unmatched_bp():
    """The reference's (unmatched, voxel-driven) cone BP is an approximate
    adjoint; the matched gradient A^T r stays close to it."""
    import torch
    from astra import _core, _lib, autograd as AG, projector3d
    pid, _, _ = _cone(48, 60, 64)
    r = _rand((64, 60, 64), 11)
    matched = AG.back_project(r, pid)
    P = projector3d.get(pid)
    unmatched = torch.empty_like(matched)
    _core.bp(P, r, unmatched, _lib.MODE_SET)
    rel = rel_l2(matched.cpu().numpy(), unmatched.cpu().numpy())
    assert rel <= 5e-2, rel


def test_back_project_gradient_is_forward_projection():
    import torch
    from astra import autograd as AG
    pid, _, _ = _cone(20, 24, 28)
    y = _rand((28, 24, 28), 12).requires_grad_(True)
    w = _rand((20, 20, 20), 13)
    (AG.back_project(y, pid) * w).sum().backward()
    ref = AG.forward_project(w, pid)
    assert rel_l2(y.grad.cpu().numpy(), ref.cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("keep", [False, True], ids=["from_checkpoint", "keep_segment"])
@pytest.mark.parametrize("lo", [None, 0.0])
def test_unrolled_sirt_recompute_equals_stored(lo, keep):
    """Gradients of an unrolled SIRT (20 iterations) w.r.t. b and x0: the
    recomputing backward (each iterate from its checkpoint, or a segment
    at a time) equals autograd through every stored iterate."""
    import torch
    from astra import autograd as AG, phantom
    n, na, nd = 32, 36, 48
    pid, _, _ = _cone(n, na, nd)
    truth = phantom.ellipsoids_3d_device(n, count=6, seed=3, device="cuda")
    b0 = AG.forward_project(truth, pid).detach()
    target = torch.roll(truth, 1, dims=2)
    W = AG.sirt_weights(pid)
    out = []
    fns = (lambda *a, **k: AG.sirt_unrolled(*a, keep_segment=keep, **k), AG.sirt_unrolled_stored)
    for fn in fns:
        b = b0.clone().requires_grad_(True)
        x0 = torch.full((n, n, n), 0.1, device="cuda").requires_grad_(True)
        x = fn(b, pid, 20, x0=x0, min_constraint=lo, weights=W)
        loss = ((x.double() - target.double()) ** 2).sum()
        gb, gx = torch.autograd.grad(loss, (b, x0))
        out.append((x.detach(), gb, gx))
    (xa, gba, gxa), (xb, gbb, gxb) = out
    # the projectors are deterministic, but the recomputing backward applies
    # the adjoint of each SIRT step in its own order of float operations,
    # while the stored run goes through autograd's per-operation backward:
    # the two agree to rounding, and after 20 iterations (and, with the
    # constraint, clamp masks of voxels sitting at 0) the iterated bar 1e-4
    # applies
    assert rel_l2(xa.cpu().numpy(), xb.cpu().numpy()) <= 1e-6
    assert rel_l2(gba.cpu().numpy(), gbb.cpu().numpy()) <= 1e-4
    assert rel_l2(gxa.cpu().numpy(), gxb.cpu().numpy()) <= 1e-4


def test_unrolled_sirt_finite_differences():
    """x_N(b) is linear in b without constraints, so L(b) = ||x_N - t||^2 is
    quadratic: central differences check the recomputed gradient."""
    import torch
    from astra import autograd as AG, phantom
    n, na, nd = 24, 30, 36
    pid, _, _ = _cone(n, na, nd)
    truth = phantom.ellipsoids_3d_device(n, count=5, seed=4, device="cuda")
    b = AG.forward_project(truth, pid).detach().requires_grad_(True)
    target = truth * 0.5
    W = AG.sirt_weights(pid)

    def L(bb):
        return ((AG.sirt_unrolled(bb, pid, 8, weights=W).double() - target.double()) ** 2).sum()

    (g,) = torch.autograd.grad(L(b), b)
    d = _rand(tuple(b.shape), 14) - 0.5
    eps = 1.0
    with torch.no_grad():
        fd = float(L(b + eps * d) - L(b - eps * d)) / (2 * eps)
    an = float((g.double() * d.double()).sum())
    assert abs(fd - an) <= 1e-3 * abs(an), (fd, an)


@pytest.mark.parametrize("ss", [1, 2])
def test_matched_bp_all_directions_anisotropic(ss):
    """Matched adjoint on a tilted cone_vec geometry whose angles fall in all
    three principal-axis groups, with a non-cubic volume of anisotropic
    voxels: against the oracle transpose, and the FP64 dot-product test."""
    import torch
    import astra
    from astra import autograd as AG
    from oracle import geometry as OG, proj3d as O3
    rng = np.random.default_rng(71)
    na, nu, nv = 18, 30, 26
    vecs = np.zeros((na, 12))
    for i in range(na):
        d = rng.normal(size=3)
        if i % 3 == 2:
            d[2] = 3.0 * np.sign(d[2] + 1e-3)
        d /= np.linalg.norm(d)
        u = np.cross(d, [0.3, 0.2, 1.0]); u /= np.linalg.norm(u)
        v = np.cross(d, u)
        vecs[i] = np.concatenate([d * 90.0, -d * 45.0, u * 1.1, v * 0.9])
    vg = astra.create_vol_geom(20, 24, 16, -15, 15, -10, 10, -6, 10)
    pg = astra.create_proj_geom("cone_vec", nv, nu, vecs)
    pid = astra.create_projector("cuda3d", pg, vg, {"DetectorSuperSampling": ss})
    ovec, vs = OG.normalise_3d(OG.conevec_from_astra(vecs, nu, nv), "cone", vg)
    assert {O3.cone_direction(ovec[a], nu, nv) for a in range(na)} == {0, 1, 2}
    y = _rand((nv, na, nu), 72)
    got = AG.back_project(y, pid).cpu().numpy()
    ref = O3.cone_fp_T(y.cpu().numpy().astype(np.float64), ovec, (16, 20, 24), vs, ss=ss)
    assert rel_l2(got, ref) <= 1e-5
    x = _rand((16, 20, 24), 73)
    lhs = float((AG.forward_project(x, pid).double() * y.double()).sum())
    rhs = float((x.double() * torch.from_numpy(got).cuda().double()).sum())
    assert abs(lhs - rhs) / abs(lhs) <= 1e-5
