"""The z-constant 3D BP with detector footprints staged in LDS
(ASTRA_AMD_BP3D_LDS=1) against the kernel that reads the taps from global
memory: same taps, weights and order of sums, so the volumes are equal bit
for bit.  Cases cover the three kernel kinds (cone fr^2 and FDK weighting,
cone 2d_weighting, par3d), LDS row layouts (odd or padded row lengths),
SET / ADD and the SIRT update, more than 256
angles per launch (two footprint groups), volumes whose edge blocks are
partly outside, and a source close enough that footprints overflow the
tile (those angles fall back to global memory).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(kind, nx, ny, nz, na, nu, nv, sod=300.0, odd=200.0):
    import astra
    vg = astra.create_vol_geom(ny, nx, nz)
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    if kind == "par3d":
        pg = astra.create_proj_geom("parallel3d", 1.1, 1.2, nv, nu, ang)
    else:
        pg = astra.create_proj_geom("cone", 1.3, 1.4, nv, nu, ang, sod, odd)
    rng = np.random.default_rng(nx * 7 + na)
    proj = rng.random((nv, na, nu), dtype=np.float32)
    return astra, vg, pg, proj


CASES = [
    ("cone", 40, 36, 28, 30, 64, 48, 300.0, 200.0),
    ("cone", 33, 47, 19, 300, 72, 40, 300.0, 200.0),    # two footprint groups, ragged blocks
    ("par3d", 45, 38, 21, 50, 64, 32, 0, 0),
    ("cone", 48, 48, 32, 24, 96, 96, 40.0, 400.0),      # near source: tiles overflow
]


@pytest.mark.parametrize("pitch", ["-1", "0", "8", "3"], ids=["odd_rows", "pm0", "pm8", "pm3"])
@pytest.mark.parametrize("case", CASES, ids=["cone", "cone_groups", "par3d", "cone_overflow"])
def test_lds_bp_equals_global_bp(case, pitch, monkeypatch):
    kind, nx, ny, nz, na, nu, nv, sod, odd = case
    astra, vg, pg, proj = _case(kind, nx, ny, nz, na, nu, nv, sod, odd)
    monkeypatch.setenv("ASTRA_AMD_BP3D_LDS_PITCH", pitch)
    out = {}
    for lds in (False, True):
        monkeypatch.setenv("ASTRA_AMD_BP3D_LDS", "1" if lds else "0")
        bid, bp = astra.create_backprojection3d_gpu(proj, pg, vg)
        astra.data3d.delete(bid)
        out[lds] = bp
    assert np.abs(out[False]).max() > 0
    assert np.array_equal(out[True], out[False])


@pytest.mark.parametrize("opt", [{}, {"ProjectionKernel": "2d_weighting"}])
def test_lds_bp_kernels_and_modes(opt, monkeypatch):
    """Projector kernels through direct BP (SET, then ADD) and FDK."""
    import astra
    import astra.experimental
    astra_, vg, pg, proj = _case("cone", 36, 30, 26, 40, 60, 44)
    cfg = {"type": "cuda3d", "ProjectionGeometry": pg, "VolumeGeometry": vg, "options": dict(opt)}
    pid = astra.projector3d.create(cfg)
    res = {}
    for lds in (False, True):
        monkeypatch.setenv("ASTRA_AMD_BP3D_LDS", "1" if lds else "0")
        v = np.zeros((26, 30, 36), np.float32)
        astra.experimental.direct_BP3D(pid, v, proj)
        once = v.copy()
        astra.experimental.direct_FPBP3D(pid, v, proj, 0, "BP")  # MODE_ADD
        f = np.zeros_like(v)
        sid = astra.data3d.link("-sino", pg, proj.copy())
        vid = astra.data3d.link("-vol", vg, f)
        astra.experimental.accumulate_FDK(pid, vid, sid)
        astra.data3d.delete([sid, vid])
        res[lds] = (once, v, f)
    for a, b in zip(res[False], res[True]):
        assert np.abs(a).max() > 0 and np.array_equal(a, b)
    astra.projector3d.delete(pid)


def test_lds_bp_sirt_update(monkeypatch):
    """The SIRT update epilogue (weights, clamp) after an LDS BP."""
    import astra
    astra_, vg, pg, proj = _case("cone", 32, 32, 24, 36, 56, 40)
    res = {}
    for lds in (False, True):
        monkeypatch.setenv("ASTRA_AMD_BP3D_LDS", "1" if lds else "0")
        sid = astra.data3d.create("-sino", pg, proj)
        rid = astra.data3d.create("-vol", vg, 0.0)
        cfg = astra.astra_dict("SIRT3D_CUDA")
        cfg.update(ProjectionDataId=sid, ReconstructionDataId=rid, option={"MinConstraint": 0.0})
        aid = astra.algorithm.create(cfg)
        astra.algorithm.run(aid, 5)
        res[lds] = astra.data3d.get(rid)
        astra.algorithm.delete(aid)
        astra.data3d.delete([sid, rid])
    assert np.abs(res[False]).max() > 0 and np.array_equal(res[True], res[False])
