"""2D projectors that serve two taps with one 16-byte load (two
neighbouring detectors of the FP, two rows of the BP; opt-in with
ASTRA_AMD_FP2D_PAIRS=1 / ASTRA_AMD_BP2D_PAIRS=1) against the default
projectors with one load per tap: same texels, same arithmetic, so the
outputs are equal bit for bit.  Cases: parallel and fan beams, odd detector counts
(a last pair of one), fan rays of one pair running along different axes,
and geometries whose pairs are more than two texels apart (wide detector
bins, a source close to the volume), which take a second load."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("parallel", 70, 37, 45, 91, 1.0, None),
    ("parallel", 64, 64, 33, 40, 2.7, None),            # bins 2.7 px apart: second loads
    ("fanflat", 83, 61, 50, 121, 1.3, (300.0, 200.0)),
    ("fanflat", 96, 96, 40, 200, 1.0, (60.0, 400.0)),    # near source
    ("fanflat", 129, 40, 77, 190, 0.8, (500.0, 300.0)),
]
IDS = ["par", "par_wide_bins", "fan", "fan_near", "fan_ragged"]


def _geoms(kind, w, h, na, nd, dw, fan):
    import astra
    vg = astra.create_vol_geom(h, w)
    ang = np.linspace(0.05, 0.05 + 2 * np.pi, na, endpoint=False)
    if kind == "parallel":
        pg = astra.create_proj_geom("parallel", dw, nd, ang)
    else:
        pg = astra.create_proj_geom("fanflat", dw, nd, ang, fan[0], fan[1])
    return vg, pg


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_fp2d_pairs_equal_single(case, monkeypatch):
    import astra
    kind, w, h, na, nd, dw, fan = case
    vg, pg = _geoms(kind, w, h, na, nd, dw, fan)
    vol = np.random.default_rng(w + nd).random((h, w), dtype=np.float32)
    out = {}
    for pairs in ("0", "1"):
        monkeypatch.setenv("ASTRA_AMD_FP2D_PAIRS", pairs)
        pid = astra.create_projector("cuda", pg, vg)
        sid, sino = astra.create_sino(vol, pid)
        astra.data2d.delete(sid)
        astra.projector.delete(pid)
        out[pairs] = sino
    assert np.abs(out["0"]).max() > 0
    assert np.array_equal(out["0"], out["1"])


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_bp2d_pairs_equal_single(case, monkeypatch):
    import astra
    kind, w, h, na, nd, dw, fan = case
    vg, pg = _geoms(kind, w, h, na, nd, dw, fan)
    sino = np.random.default_rng(h + na).random((na, nd), dtype=np.float32)
    out = {}
    for pairs in ("0", "1"):
        monkeypatch.setenv("ASTRA_AMD_BP2D_PAIRS", pairs)
        monkeypatch.setenv("ASTRA_AMD_FP2D_PAIRS", pairs)
        pid = astra.create_projector("cuda", pg, vg)
        bid, bp = astra.create_backprojection(sino, pid)
        astra.data2d.delete(bid)
        res = [bp]
        algs = ["SIRT_CUDA"] + (["FBP_CUDA"] if kind == "fanflat" else [])
        for alg in algs:
            sid = astra.data2d.create("-sino", pg, sino)
            rid = astra.data2d.create("-vol", vg, 0)
            cfg = astra.astra_dict(alg)
            cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
            aid = astra.algorithm.create(cfg)
            astra.algorithm.run(aid, 3 if alg == "SIRT_CUDA" else 1)
            res.append(astra.data2d.get(rid))
            astra.algorithm.delete(aid)
            astra.data2d.delete([sid, rid])
        astra.projector.delete(pid)
        out[pairs] = res
    for a, b in zip(out["0"], out["1"]):
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b)
