"""The 3D FP with volume slices staged in LDS for groups of 4 angles
(ASTRA_AMD_FP3D_LDS=1) against the FP that reads its taps from global
memory: each ray takes the same samples in the same order, so the
projections are equal bit for bit.  Cases cover cone and par3d, detectors
whose tiles are ragged or one pixel wide, angle counts that do not fill the
last group, all three principal axes (cone_vec), a source close enough that
the staged rectangles overflow (those chunks read global memory), and the
SIRT residual epilogue."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("cone", 40, 36, 28, 30, 64, 48, 300.0, 200.0),
    ("cone", 33, 47, 19, 37, 72, 40, 300.0, 200.0),
    ("par3d", 45, 38, 21, 50, 64, 32, 0, 0),
    ("cone", 20, 20, 20, 9, 1, 23, 100.0, 50.0),      # one detector column
    ("cone", 48, 48, 32, 24, 96, 96, 40.0, 400.0),    # near source: rectangles overflow
]


def _geom(kind, nx, ny, nz, na, nu, nv, sod, odd):
    import astra
    vg = astra.create_vol_geom(ny, nx, nz)
    ang = np.linspace(0.1, 0.1 + 2 * np.pi, na, endpoint=False)
    if kind == "par3d":
        pg = astra.create_proj_geom("parallel3d", 1.1, 1.2, nv, nu, ang)
    else:
        pg = astra.create_proj_geom("cone", 1.3, 1.4, nv, nu, ang, sod, odd)
    return vg, pg


def _fp_both(vg, pg, vol, monkeypatch):
    import astra
    out = {}
    for lds in ("0", "1"):
        monkeypatch.setenv("ASTRA_AMD_FP3D_LDS", lds)
        sid, p = astra.create_sino3d_gpu(vol, pg, vg)
        astra.data3d.delete(sid)
        out[lds] = p
    return out


@pytest.mark.parametrize("case", CASES, ids=["cone", "cone_ragged", "par3d", "one_column", "overflow"])
def test_lds_fp_equals_global_fp(case, monkeypatch):
    kind, nx, ny, nz, na, nu, nv, sod, odd = case
    vg, pg = _geom(kind, nx, ny, nz, na, nu, nv, sod, odd)
    rng = np.random.default_rng(nx * 5 + na)
    vol = rng.random((nz, ny, nx), dtype=np.float32)
    out = _fp_both(vg, pg, vol, monkeypatch)
    assert np.abs(out["0"]).max() > 0
    assert np.array_equal(out["0"], out["1"])


def test_lds_fp_all_principal_axes(monkeypatch):
    """cone_vec with sources along x, y and z (three direction groups)."""
    import astra
    nx, ny, nz, nu, nv = 30, 26, 22, 40, 36
    vg = astra.create_vol_geom(ny, nx, nz)
    vecs = []
    for k, src in enumerate([(200, 0, 0), (0, -180, 0), (0, 0, 220), (150, 140, 10),
                             (-10, 20, -190), (-200, 5, 3), (7, 210, -4)]):
        s = np.array(src, float)
        d = -0.5 * s
        axis = np.argmax(np.abs(s))
        u = np.zeros(3)
        u[(axis + 1) % 3] = 1.0
        v = np.cross(s / np.linalg.norm(s), u)
        v /= np.linalg.norm(v)
        vecs.append(np.concatenate([s, d, 1.1 * u, 1.2 * v]))
    pg = astra.create_proj_geom("cone_vec", nv, nu, np.array(vecs))
    rng = np.random.default_rng(3)
    vol = rng.random((nz, ny, nx), dtype=np.float32)
    out = _fp_both(vg, pg, vol, monkeypatch)
    assert np.abs(out["0"]).max() > 0
    assert np.array_equal(out["0"], out["1"])


def test_lds_fp_sirt3d(monkeypatch):
    """SIRT3D's residual FP (b - A x) * w with the staged FP."""
    import astra
    vg, pg = _geom("cone", 32, 30, 24, 36, 56, 40, 300.0, 200.0)
    rng = np.random.default_rng(9)
    proj = rng.random((40, 36, 56), dtype=np.float32)
    res = {}
    for lds in ("0", "1"):
        monkeypatch.setenv("ASTRA_AMD_FP3D_LDS", lds)
        sid = astra.data3d.create("-sino", pg, proj)
        rid = astra.data3d.create("-vol", vg, 0.0)
        cfg = astra.astra_dict("SIRT3D_CUDA")
        cfg.update(ProjectionDataId=sid, ReconstructionDataId=rid)
        aid = astra.algorithm.create(cfg)
        astra.algorithm.run(aid, 4)
        res[lds] = astra.data3d.get(rid)
        astra.algorithm.delete(aid)
        astra.data3d.delete([sid, rid])
    assert np.abs(res["0"]).max() > 0 and np.array_equal(res["0"], res["1"])
