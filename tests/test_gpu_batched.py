"""Batched 2D SIRT on the GPU: equal to per-slice SIRT_CUDA, to the FP64
oracle, and graph replays equal ungraphed runs."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

SLICES = [dict(n=20, na=17, nd=30, kind="parallel"), dict(n=33, na=41, nd=50, kind="fanflat"),
          dict(n=48, na=23, nd=72, kind="parallel"), dict(n=64, na=60, nd=96, kind="fanflat"),
          dict(n=31, na=90, nd=47, kind="fanflat"), dict(n=70, na=35, nd=105, kind="parallel")]


@pytest.fixture(scope="module")
def batch():
    import torch
    import astra
    from astra import batched, phantom
    pgs, vgs = batched.geometries(SLICES)
    B = batched.BatchedSIRT2D(pgs, vgs)
    vols = [phantom.ellipses_2d(s["n"], seed=100 + i) for i, s in enumerate(SLICES)]
    truth = B.pack_volumes(vols)
    sinos = B.forward(truth)
    torch.cuda.synchronize()
    return dict(astra=astra, B=B, pgs=pgs, vgs=vgs, vols=vols, sinos=sinos)


def test_batched_fp_matches_single_slice(batch):
    astra = batch["astra"]
    B = batch["B"]
    for i, (pg, vg) in enumerate(zip(batch["pgs"], batch["vgs"])):
        pid = astra.create_projector("cuda", pg, vg)
        _, ref = astra.create_sino(batch["vols"][i], pid)
        got = B.sinogram(batch["sinos"], i).cpu().numpy()
        assert rel_l2(got, ref) <= 1e-6, i
        astra.projector.delete(pid)


def test_batched_sirt_matches_per_slice_sirt(batch):
    astra = batch["astra"]
    B = batch["B"]
    x = B.empty_volumes()
    B.run(batch["sinos"], x, 10, use_graph=False)
    for i, (pg, vg) in enumerate(zip(batch["pgs"], batch["vgs"])):
        pid = astra.create_projector("cuda", pg, vg)
        sid = astra.data2d.create("-sino", pg, B.sinogram(batch["sinos"], i).cpu().numpy())
        rid = astra.data2d.create("-vol", vg, 0)
        cfg = astra.astra_dict("SIRT_CUDA")
        cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
        aid = astra.algorithm.create(cfg)
        astra.algorithm.run(aid, 10)
        ref = astra.data2d.get(rid)
        got = B.volume(x, i).cpu().numpy()
        assert rel_l2(got, ref) <= 1e-5, (i, rel_l2(got, ref))
        astra.algorithm.delete(aid)
        astra.data2d.delete([sid, rid])
        astra.projector.delete(pid)


def test_batched_sirt_matches_oracle(batch):
    from oracle import algorithms as OA, geometry as OG, proj2d as O2
    B = batch["B"]
    x = B.empty_volumes()
    B.run(batch["sinos"], x, 8, use_graph=True)
    i = 3  # fan slice
    s = SLICES[i]
    ang = np.linspace(0, 2 * np.pi, s["na"], endpoint=False)
    vg = batch["vgs"][i]
    v, osc = OG.normalise_2d(OG.gen_fan_projections(ang, s["nd"], 2.0 * s["n"], 1.0 * s["n"], 1.0),
                             "fan", vg)
    fp = lambda z: O2.fan_fp(z, v, s["nd"], osc)
    bp = lambda y: O2.fan_bp(y, v, (s["n"], s["n"]), osc)
    b = B.sinogram(batch["sinos"], i).cpu().numpy().astype(np.float64)
    ref = OA.sirt(fp, bp, b, np.zeros((s["n"], s["n"])), 8)
    assert rel_l2(B.volume(x, i).cpu().numpy(), ref) <= 1e-4


def test_graph_replay_equals_ungraphed(batch):
    import torch
    B = batch["B"]
    a = B.empty_volumes()
    B.run(batch["sinos"], a, 5, use_graph=True, min_constraint=0.0)
    B.run(batch["sinos"], a, 5, use_graph=True, min_constraint=0.0)  # replay of the same graph
    torch.cuda.synchronize()
    c = B.empty_volumes()
    B.run(batch["sinos"], c, 10, use_graph=False, min_constraint=0.0)
    assert torch.equal(a, c)
    assert float(a.min()) >= 0.0
