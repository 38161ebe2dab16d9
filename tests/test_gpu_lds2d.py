"""The 2D projectors with LDS-staged footprints against the same projectors
reading their taps from global memory: same taps, weights and order of
sums, so the outputs are equal bit for bit.

* single-slice BP (parallel, fan, FBP-weighted fan), ragged tiles, angle
  counts that are not a multiple of the staged block, and a source close
  enough that footprints overflow the staged row (those angles fall back);
* batched SIRT with every FP tiling (three global-memory ones with one or
  two detectors per thread, and the LDS-staged one) and every BP tiling
  (three rows-per-thread choices, each global, LDS-staged or with two rows
  per load).  The FP value of a ray does not depend on the
  tiling, so every FP tiling gives the same volumes.  The BP steps a
  pixel's coordinate from its thread's first row (as the reference steps
  it along the row loop), so the last bits depend on the rows per thread:
  the LDS-staged BP equals the global one of the same rows per thread, and
  different rows per thread agree to float rounding.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BP_CASES = [
    ("parallel", 70, 37, 45, 90, None),
    ("parallel", 64, 64, 16, 100, None),
    ("fanflat", 83, 61, 33, 120, (300.0, 200.0)),
    ("fanflat", 96, 96, 50, 200, (60.0, 400.0)),   # near source: wide footprints
    ("fanflat", 129, 40, 721, 190, (500.0, 300.0)),
]


def _geoms(kind, w, h, na, nd, fan):
    import astra
    vg = astra.create_vol_geom(h, w)
    ang = np.linspace(0.2, 0.2 + 2 * np.pi, na, endpoint=False)
    if kind == "parallel":
        pg = astra.create_proj_geom("parallel", 1.1, nd, ang)
    else:
        pg = astra.create_proj_geom("fanflat", 1.3, nd, ang, fan[0], fan[1])
    return vg, pg


@pytest.mark.parametrize("case", BP_CASES, ids=["par_ragged", "par_16", "fan", "fan_near", "fan_721"])
def test_bp2d_lds_equals_global(case, monkeypatch):
    import astra
    kind, w, h, na, nd, fan = case
    vg, pg = _geoms(kind, w, h, na, nd, fan)
    rng = np.random.default_rng(w + na)
    sino = rng.random((na, nd), dtype=np.float32)
    out = {}
    for lds in ("0", "1"):
        monkeypatch.setenv("ASTRA_AMD_BP2D_LDS", lds)
        pid = astra.create_projector("cuda", pg, vg)
        bid, bp = astra.create_backprojection(sino, pid)
        astra.data2d.delete(bid)
        res = [bp]
        if kind == "fanflat":
            # FBP: the FBP-weighted BP after the fan pre-weighting and filter
            sid = astra.data2d.create("-sino", pg, sino)
            rid = astra.data2d.create("-vol", vg, 0)
            cfg = astra.astra_dict("FBP_CUDA")
            cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
            aid = astra.algorithm.create(cfg)
            astra.algorithm.run(aid)
            res.append(astra.data2d.get(rid))
            astra.algorithm.delete(aid)
            astra.data2d.delete([sid, rid])
        # SIRT: the update epilogue after the staged BP
        sid = astra.data2d.create("-sino", pg, sino)
        rid = astra.data2d.create("-vol", vg, 0)
        cfg = astra.astra_dict("SIRT_CUDA")
        cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid,
                   option={"MinConstraint": 0.0})
        aid = astra.algorithm.create(cfg)
        astra.algorithm.run(aid, 3)
        res.append(astra.data2d.get(rid))
        astra.algorithm.delete(aid)
        astra.data2d.delete([sid, rid])
        astra.projector.delete(pid)
        out[lds] = res
    for a, b in zip(out["0"], out["1"]):
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b)


SLICES = [dict(n=20, na=17, nd=30, kind="parallel"), dict(n=33, na=41, nd=50, kind="fanflat"),
          dict(n=48, na=23, nd=72, kind="parallel"), dict(n=64, na=60, nd=96, kind="fanflat"),
          dict(n=131, na=90, nd=197, kind="fanflat"), dict(n=70, na=35, nd=105, kind="parallel"),
          dict(n=200, na=47, nd=300, kind="parallel")]


def _batched_sirt(fp_cfg, bp_cfg, monkeypatch, iters=5):
    import torch
    from astra import batched, phantom
    monkeypatch.setenv("ASTRA_AMD_BATCH_FP_CFG", str(fp_cfg))
    monkeypatch.setenv("ASTRA_AMD_BATCH_BP_CFG", str(bp_cfg))
    pgs, vgs = batched.geometries(SLICES)
    B = batched.BatchedSIRT2D(pgs, vgs)
    vols = [phantom.ellipses_2d(s["n"], seed=300 + i) for i, s in enumerate(SLICES)]
    sinos = B.forward(B.pack_volumes(vols))
    x = B.empty_volumes()
    B.run(sinos, x, iters, use_graph=False)
    torch.cuda.synchronize()
    out = x.cpu().numpy().copy()
    B.release()
    return out


@pytest.mark.parametrize("rows", [0, 1, 2], ids=["py4", "py2", "py8"])
def test_batched_tilings(rows, monkeypatch):
    ref = _batched_sirt(0, rows, monkeypatch)
    assert np.abs(ref).max() > 0
    # every FP tiling (one or two detectors per thread, global or LDS), with
    # the global, LDS-staged and row-pair BP of these rows per thread
    for fp_cfg in (0, 1, 2, 3, 4, 5, 6):
        for bp_cfg in (rows, rows + 3, rows + 6):
            if (fp_cfg, bp_cfg) == (0, rows):
                continue
            got = _batched_sirt(fp_cfg, bp_cfg, monkeypatch)
            assert np.array_equal(got, ref), (fp_cfg, bp_cfg, float(np.abs(got - ref).max()))
    # other rows per thread: equal to float rounding
    other = _batched_sirt(0, (rows + 1) % 3, monkeypatch)
    err = np.linalg.norm(other - ref) / np.linalg.norm(ref)
    assert err <= 1e-6, err
