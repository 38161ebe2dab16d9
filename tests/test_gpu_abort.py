"""Stopping an iterative run between iterations (the reference's
setShouldAbortHook / shouldAbort, include/astra/Globals.h:156-164, asked
before every iteration, src/CudaSirtAlgorithm.cpp:347).

The native hook is asked once per iteration; the run stops with the iterate
reached, so a run of 10 stopped before iteration k equals a plain run of k
iterations bit for bit (same kernels, same order).  Through the Python API
the hook reads a flag that SIGINT sets: Ctrl-C raises KeyboardInterrupt with
the partial reconstruction written back.
"""
import os
import signal

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(kind):
    import astra
    if kind == "sirt2d":
        vg = astra.create_vol_geom(48, 48)
        pg = astra.create_proj_geom("fanflat", 1.0, 64, np.linspace(0, 2 * np.pi, 40, False),
                                    200.0, 100.0)
        mod, algo, ptype = astra.data2d, "SIRT_CUDA", "cuda"
        shape = (48, 48)
    else:
        vg = astra.create_vol_geom(24, 24, 24)
        pg = astra.create_proj_geom("cone", 1.0, 1.0, 24, 32, np.linspace(0, 2 * np.pi, 30, False),
                                    150.0, 50.0)
        mod, algo, ptype = astra.data3d, "CGLS3D_CUDA", "cuda3d"
        shape = (24, 24, 24)
    rng = np.random.default_rng(3)
    vol = rng.random(shape, dtype=np.float32)
    pid = astra.create_projector(ptype, pg, vg)
    if ptype == "cuda":
        sid, _ = astra.create_sino(vol, pid)
    else:
        sid, _ = astra.create_sino3d_gpu(vol, pg, vg)
    return astra, mod, algo, pid, vg, sid


def _run(astra, mod, algo, pid, vg, sid, iters):
    rid = mod.create("-vol", vg, 0.0)
    cfg = astra.astra_dict(algo)
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
    aid = astra.algorithm.create(cfg)
    try:
        astra.algorithm.run(aid, iters)
        return None, mod.get(rid)
    except KeyboardInterrupt as e:
        return e, mod.get(rid)
    finally:
        astra.algorithm.delete(aid)
        mod.delete(rid)


@pytest.fixture
def quiet_sigint():
    """A no-op SIGINT handler around the test, so that a signal handled
    after a run has ended cannot interrupt the test session."""
    prev = signal.signal(signal.SIGINT, lambda signum, frame: None)
    yield
    signal.signal(signal.SIGINT, prev)


@pytest.mark.parametrize("kind", ["sirt2d", "cgls3d"])
def test_hook_stops_before_iteration_k(kind, monkeypatch, quiet_sigint):
    from astra import algorithm
    astra, mod, algo, pid, vg, sid = _problem(kind)
    calls = []

    def stop_at_5th():
        calls.append(1)
        return len(calls) >= 5

    monkeypatch.setattr(algorithm, "_should_abort", stop_at_5th)
    err, got = _run(astra, mod, algo, pid, vg, sid, 10)
    assert isinstance(err, KeyboardInterrupt) and "after 4 iterations" in str(err)
    assert len(calls) == 5
    monkeypatch.undo()
    none, want = _run(astra, mod, algo, pid, vg, sid, 4)
    assert none is None and np.abs(want).max() > 0
    assert np.array_equal(got, want)
    # the hook is removed after the run: a full run is not stopped
    none, full = _run(astra, mod, algo, pid, vg, sid, 10)
    assert none is None and not np.array_equal(full, want)


def test_sigint_stops_a_run(monkeypatch, quiet_sigint):
    """SIGINT sent from inside the run (at the 3rd hook call) stops it at
    the next iteration boundary the handler has been seen at."""
    from astra import algorithm
    astra, mod, algo, pid, vg, sid = _problem("sirt2d")
    calls = []
    stopped = []

    def kill_at_3rd():
        calls.append(1)
        if len(calls) == 3:
            os.kill(os.getpid(), signal.SIGINT)
        if algorithm._stop.is_set():
            stopped.append(len(calls))
            return True
        return False

    monkeypatch.setattr(algorithm, "_should_abort", kill_at_3rd)
    err, got = _run(astra, mod, algo, pid, vg, sid, 50)
    assert isinstance(err, KeyboardInterrupt)
    assert stopped and 3 <= stopped[0] <= 5
    monkeypatch.undo()
    none, want = _run(astra, mod, algo, pid, vg, sid, stopped[0] - 1)
    assert none is None and np.array_equal(got, want)
