"""The GPU 2D parallel path against ASTRA's own CPU projectors, with the
error stated.  The expected outputs are the reference's CParallelBeamLinear-
KernelProjector2D and CParallelBeamStripKernelProjector2D (FP, BP) and 10
iterations of its CPU SIRT, built from the reference sources
(tests/golden/make_golden.py, make_strip_golden.py): a 64 x 64 image of
seeded ellipses, 45 angles over [0, pi), 96 detectors of width 1.0.

Stated differences (rel-L2), each asserted below with headroom:
* GPU FP vs CPU linear FP: same discretisation (linear interpolation at
  row / column centres), 2.6e-6 in FP64 — held to 1e-5;
* GPU matched BP (exact transpose of the GPU FP) vs CPU linear BP (exact
  transpose of the CPU FP): 4.7e-6 in FP64 — held to 1e-5;
* GPU FP vs CPU strip FP (area-weighted): 1.6e-3 — held to 5e-3;
* GPU BP (voxel-driven, the reference GPU BP's model) vs CPU linear BP:
  1.4e-2, vs CPU strip BP: 2.6e-3 — held to 3e-2 and 6e-3;
* GPU SIRT (10 iterations, voxel-driven BP) vs CPU linear SIRT: 7.8e-3
  in FP64 — held to 2e-2.
"""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def case():
    import astra
    lin = np.load(os.path.join(GOLD, "linear_cpu_64.npz"))
    strip = np.load(os.path.join(GOLD, "strip_cpu_64.npz"))
    vg = astra.create_vol_geom(64, 64)
    pg = astra.create_proj_geom("parallel", float(lin["det_width"]), int(lin["n_det"]),
                                lin["angles"].astype(np.float64))
    pid = astra.create_projector("cuda", pg, vg)
    yield astra, pid, vg, pg, lin, strip
    astra.projector.delete(pid)


def test_fp_vs_cpu_linear_and_strip(case):
    astra, pid, vg, pg, lin, strip = case
    sid, fp = astra.create_sino(lin["vol"], pid)
    astra.data2d.delete(sid)
    e_lin, e_strip = rel_l2(fp, lin["fp"]), rel_l2(fp, strip["fp"])
    print("GPU FP vs CPU linear %.3g, vs CPU strip %.3g" % (e_lin, e_strip))
    assert e_lin <= 1e-5 and e_strip <= 5e-3


def test_matched_bp_vs_cpu_linear_bp(case):
    import torch
    from astra import autograd
    astra, pid, vg, pg, lin, strip = case
    s = torch.from_numpy(lin["sino_in"].astype(np.float32)).cuda()
    bp = autograd.back_project(s, pid).cpu().numpy()
    e = rel_l2(bp, lin["bp"])
    print("GPU matched BP vs CPU linear BP %.3g" % e)
    assert e <= 1e-5


def test_default_bp_vs_cpu_bps(case):
    astra, pid, vg, pg, lin, strip = case
    bid, bp = astra.create_backprojection(lin["sino_in"], pid)
    astra.data2d.delete(bid)
    e_lin, e_strip = rel_l2(bp, lin["bp"]), rel_l2(bp, strip["bp"])
    print("GPU BP vs CPU linear BP %.3g, vs CPU strip BP %.3g" % (e_lin, e_strip))
    assert e_lin <= 3e-2 and e_strip <= 6e-3


def test_sirt_vs_cpu_linear_sirt(case):
    astra, pid, vg, pg, lin, strip = case
    sid = astra.data2d.create("-sino", pg, lin["fp"])
    rid = astra.data2d.create("-vol", vg, 0.0)
    cfg = astra.astra_dict("SIRT_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
    aid = astra.algorithm.create(cfg)
    astra.algorithm.run(aid, 10)
    got = astra.data2d.get(rid)
    e = rel_l2(got, lin["sirt10"])
    print("GPU SIRT (10 its) vs CPU linear SIRT %.3g" % e)
    assert e <= 2e-2
    astra.algorithm.delete(aid)
    astra.data2d.delete([sid, rid])
