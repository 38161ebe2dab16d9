"""The 3D FP's workgroup map: each XCD takes an eighth of a direction
group's angles, cut into equal blocks of at most ASTRA_AMD_FP3D_ANGLE_BLOCK
angles (default 64).  Which workgroup computes a ray does not change the
ray's arithmetic, so every block size gives the same projections bit for
bit, and the default is checked against the oracle.  Cases: fewer angles
than XCDs, counts that leave XCD ranges and blocks uneven, par3d with
ragged tiles, all three principal axes (cone_vec) and detector
supersampling."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fp(pid, vol, shape, monkeypatch, block):
    import torch
    from astra import experimental
    monkeypatch.setenv("ASTRA_AMD_FP3D_ANGLE_BLOCK", block)
    proj = torch.empty(shape, dtype=torch.float32, device="cuda")
    experimental.direct_FP3D(pid, torch.from_numpy(vol).cuda(), proj)
    return proj.cpu().numpy()


def _check(pg, vg, shape, monkeypatch, opts=None, seed=0):
    import astra
    pid = astra.create_projector("cuda3d", pg, vg, opts or {})
    rng = np.random.default_rng(seed)
    nz, ny, nx = vg["GridSliceCount"], vg["GridRowCount"], vg["GridColCount"]
    vol = rng.random((nz, ny, nx), dtype=np.float32)
    ref = _fp(pid, vol, shape, monkeypatch, "64")
    assert np.abs(ref).max() > 0
    for block in ("1", "5", "13"):
        assert np.array_equal(ref, _fp(pid, vol, shape, monkeypatch, block)), block
    astra.projector3d.delete(pid)
    return vol, ref


@pytest.mark.parametrize("na", [2, 7, 97])
def test_fp_map_cone(na, monkeypatch):
    import astra
    from oracle import geometry as OG, proj3d as O3
    n, nd = 24, 36
    vg = astra.create_vol_geom(n, n, n)
    ang = np.linspace(0.05, 0.05 + 2 * np.pi, na, endpoint=False)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, ang, 80.0, 40.0)
    vol, ref = _check(pg, vg, (nd, na, nd), monkeypatch, seed=na)
    v, vs = OG.normalise_3d(OG.gen_cone_projections(ang, nd, nd, 80.0, 40.0, 1.0, 1.0), "cone",
                            OG.vol_geom_3d(n, n, n))
    want = O3.cone_fp(vol.astype(np.float64), v, nd, nd, vs)
    assert np.linalg.norm(ref - want) <= 1e-5 * np.linalg.norm(want)


def test_fp_map_par3d_ragged(monkeypatch):
    import astra
    vg = astra.create_vol_geom(37, 41, 29)
    ang = np.linspace(0, np.pi, 61, endpoint=False)
    pg = astra.create_proj_geom("parallel3d", 1.1, 0.9, 35, 53, ang)
    _check(pg, vg, (35, 61, 53), monkeypatch, seed=4)


def test_fp_map_all_axes(monkeypatch):
    import astra
    nx, ny, nz, nu, nv = 30, 26, 22, 40, 36
    vg = astra.create_vol_geom(ny, nx, nz)
    vecs = []
    rng = np.random.default_rng(11)
    for k in range(40):
        s = rng.normal(size=3)
        s *= 200.0 / np.linalg.norm(s)
        axis = np.argmax(np.abs(s))
        u = np.zeros(3)
        u[(axis + 1) % 3] = 1.0
        u -= s * (s @ u) / (s @ s)
        u /= np.linalg.norm(u)
        v = np.cross(s / np.linalg.norm(s), u)
        vecs.append(np.concatenate([s, -0.5 * s, 1.1 * u, 1.2 * v]))
    pg = astra.create_proj_geom("cone_vec", nv, nu, np.array(vecs))
    _check(pg, vg, (nv, 40, nu), monkeypatch, seed=5)


def test_fp_map_supersampled(monkeypatch):
    import astra
    n, na, nd = 24, 30, 32
    vg = astra.create_vol_geom(n, n, n)
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, ang, 90.0, 45.0)
    _check(pg, vg, (nd, na, nd), monkeypatch, {"DetectorSuperSampling": 2}, seed=6)
