"""The batched 2D SIRT as one persistent launch over device work queues:
the same floats as the per-stage launches (graphed or not) for every tiling
it supports, across consecutive runs and constraint changes, on slices of
uneven size, angle count and geometry; a wait past a tiny spin limit stops
the run with an error and leaves the next run correct."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SLICES = [dict(n=20, na=17, nd=30, kind="parallel"), dict(n=33, na=41, nd=50, kind="fanflat"),
          dict(n=48, na=23, nd=72, kind="parallel"), dict(n=64, na=60, nd=96, kind="fanflat"),
          dict(n=31, na=90, nd=47, kind="fanflat"), dict(n=70, na=35, nd=105, kind="parallel"),
          dict(n=150, na=97, nd=225, kind="parallel"), dict(n=203, na=131, nd=305, kind="fanflat"),
          dict(n=1, na=1, nd=1, kind="parallel"), dict(n=129, na=3, nd=194, kind="fanflat")]


def _batch(slices, seed0=100):
    import torch
    from astra import batched, phantom
    pgs, vgs = batched.geometries(slices)
    B = batched.BatchedSIRT2D(pgs, vgs)
    vols = [phantom.ellipses_2d(s["n"], seed=seed0 + i) for i, s in enumerate(slices)]
    truth = B.pack_volumes(vols)
    sinos = B.forward(truth)
    torch.cuda.synchronize()
    return B, sinos


@pytest.fixture(scope="module")
def batch():
    return _batch(SLICES)


def test_persistent_equals_stage_launches(batch):
    import torch
    B, sinos = batch
    a = B.empty_volumes()
    B.run(sinos, a, 7, min_constraint=0.0, persistent=True)
    b = B.empty_volumes()
    B.run(sinos, b, 7, min_constraint=0.0, use_graph=False)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert float(a.abs().max()) > 0


def test_persistent_consecutive_runs(batch):
    """run(3); run(4) == run(7), replays of the captured run included, and a
    change of constraints re-captures it."""
    import torch
    B, sinos = batch
    a = B.empty_volumes()
    B.run(sinos, a, 3, persistent=True)
    B.run(sinos, a, 4, persistent=True)
    c = B.empty_volumes()
    B.run(sinos, c, 4, persistent=True)
    B.run(sinos, c, 3, persistent=True)
    d = B.empty_volumes()
    B.run(sinos, d, 7, use_graph=True)
    torch.cuda.synchronize()
    assert torch.equal(a, d)
    assert torch.equal(c, d)
    e = B.empty_volumes()
    B.run(sinos, e, 4, min_constraint=0.0, max_constraint=0.5, persistent=True)
    f = B.empty_volumes()
    B.run(sinos, f, 4, min_constraint=0.0, max_constraint=0.5, use_graph=False)
    torch.cuda.synchronize()
    assert torch.equal(e, f)
    assert float(e.max()) <= 0.5 and float(e.min()) >= 0.0


@pytest.mark.parametrize("fp_cfg", [0, 1, 2])
@pytest.mark.parametrize("bp_cfg", [0, 2, 3, 5])
def test_persistent_every_tiling(fp_cfg, bp_cfg, monkeypatch):
    import torch
    monkeypatch.setenv("ASTRA_AMD_BATCH_FP_CFG", str(fp_cfg))
    monkeypatch.setenv("ASTRA_AMD_BATCH_BP_CFG", str(bp_cfg))
    B, sinos = _batch(SLICES[:8], seed0=7)
    a = B.empty_volumes()
    B.run(sinos, a, 4, min_constraint=0.0, persistent=True)
    b = B.empty_volumes()
    B.run(sinos, b, 4, min_constraint=0.0, use_graph=False)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    B.release()


@pytest.mark.parametrize("fp_cfg,bp_cfg", [(4, 3), (5, 6), (3, 8), (6, 5)])
def test_persistent_with_tilings_it_maps(fp_cfg, bp_cfg, monkeypatch):
    """Chosen tilings the persistent kernel has no form of (LDS and pair
    FP, row-pair BP) run with the plain tiling of the same detectors /
    rows per thread: the same floats."""
    import torch
    monkeypatch.setenv("ASTRA_AMD_BATCH_FP_CFG", str(fp_cfg))
    monkeypatch.setenv("ASTRA_AMD_BATCH_BP_CFG", str(bp_cfg))
    B, sinos = _batch(SLICES[:8], seed0=9)
    a = B.empty_volumes()
    B.run(sinos, a, 3, min_constraint=0.0, persistent=True)
    b = B.empty_volumes()
    B.run(sinos, b, 3, min_constraint=0.0, use_graph=False)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    B.release()


@pytest.mark.parametrize("chunk", ["1", "2", "3"])
def test_persistent_in_consecutive_launches(chunk, batch, monkeypatch):
    """A run split into persistent launches of `chunk` iterations (as runs
    whose dequeue counters would pass 2^31 are) equals one launch."""
    import torch
    B, sinos = batch
    ref = B.empty_volumes()
    B.run(sinos, ref, 7, min_constraint=0.0, persistent=True)
    monkeypatch.setenv("ASTRA_AMD_BATCH_PERSIST_ITERS", chunk)
    x = B.empty_volumes()
    B.run(sinos, x, 7, min_constraint=0.0, persistent=True)
    torch.cuda.synchronize()
    assert torch.equal(x, ref)


def test_persistent_register_limited_variant(batch, monkeypatch):
    import torch
    B, sinos = batch
    monkeypatch.setenv("ASTRA_AMD_BATCH_PERSIST_WPE", "6")
    a = B.empty_volumes()
    B.run(sinos, a, 5, persistent=True)
    b = B.empty_volumes()
    B.run(sinos, b, 5, use_graph=False)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_persistent_spin_limit_stops_and_recovers(monkeypatch):
    """With no polls allowed, the first item whose previous stage is not yet
    complete stops the run (every kernel still exits normally); the next run
    with the default limit is correct."""
    import torch
    from astra import batched
    from astra.log import AstraError
    slices = batched.config5_slices(64, seed=3, n_min=64, n_max=160, a_min=30, a_max=90)
    B, sinos = _batch(slices, seed0=50)
    ref = B.empty_volumes()
    B.run(sinos, ref, 3, use_graph=False)
    monkeypatch.setenv("ASTRA_AMD_BATCH_SPIN_LIMIT", "0")
    x = B.empty_volumes()
    stopped = False
    try:
        B.run(sinos, x, 3, persistent=True)
    except AstraError as e:
        assert "spin limit" in str(e)
        stopped = True
    if not stopped:  # every wait happened to be satisfied at its first poll
        assert torch.equal(x, ref)
    monkeypatch.delenv("ASTRA_AMD_BATCH_SPIN_LIMIT")
    y = B.empty_volumes()
    B.run(sinos, y, 3, persistent=True)
    torch.cuda.synchronize()
    assert torch.equal(y, ref)
    B.release()


def test_persistent_4096_slices():
    """The benchmarked 4096-slice batch (128-1024 px, 90-720 angles,
    parallel / fan): 3 iterations as one persistent launch equal the graphed
    per-stage launches bit for bit."""
    import torch
    from astra import batched, phantom
    slices = batched.config5_slices(4096, seed=0)
    pgs, vgs = batched.geometries(slices)
    B = batched.BatchedSIRT2D(pgs, vgs)
    truth = B.empty_volumes()
    for i, s in enumerate(slices):
        B.volume(truth, i).copy_(phantom.ellipses_2d_device(s["n"], seed=1000 + i, device="cuda"))
    sinos = B.forward(truth)
    a = B.empty_volumes()
    B.run(sinos, a, 3, min_constraint=0.0, persistent=True)
    b = B.empty_volumes()
    B.run(sinos, b, 3, min_constraint=0.0)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert float(a.max()) > 0
    B.release()
    del truth, sinos, a, b
    torch.cuda.empty_cache()
