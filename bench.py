#!/usr/bin/env python
"""Benchmark of the MI355X ASTRA projection hot path.

Headline (one JSON line on stdout): cone-beam FP+BP throughput in
Gupdates/s on the 512^3 workload -- 512^3 fp32 volume of 20 seeded
ellipsoids, 720 projections over [0, 2pi), 768 x 768 detector with pixel 1,
source-origin 1000, origin-detector 500.  One step = one forward projection
plus one back projection (both MODE_SET) with volume and projections already
resident in HBM.  Updates = FP ray-slice samples (every slice of every ray,
as the reference samples them) + BP voxel-angle updates.

Also reported: FDK seconds at the same size, per-kernel event timings of
the hot kernels (roofline of the dominant one), and the reference CPU path
(2D parallel-beam SIRT, 'linear' projector) timed on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "astra-toolbox_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """nproc, the CPUs this process may run on, and the CPU model (lscpu)."""
    import subprocess

    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "affinity_cpus": allowed, "model": model}


def cpu_baseline(iters, runs=5):
    """Reference CPU SIRT (256^2, 180 angles, 384 detectors, 'linear'),
    single thread: one warm-up run, then the median of `runs` timed runs of
    `iters` iterations (phantom, sinogram and geometry set up outside the
    timed region)."""
    from astra import phantom
    from oracle import cpu

    lib = cpu.reference()
    kind = "reference"
    if lib is None:
        lib, kind = cpu.port(), "port"
    n, na, nd = 256, 180, 384
    ang = np.linspace(0, np.pi, na, endpoint=False)
    vol = phantom.ellipses_2d(n, seed=0)
    sino = lib.fp(vol, ang, nd)
    lib.sirt(sino, ang, (n, n), 2)  # warm-up
    secs = []
    for _ in range(runs):
        t0 = time.perf_counter()
        lib.sirt(sino, ang, (n, n), iters)
        secs.append(time.perf_counter() - t0)
    dt = float(np.median(secs))
    updates = iters * 2.0 * na * nd * n  # FP + BP ray-slice samples per iteration
    return {
        "value": updates / dt / 1e9,
        "unit": "Gupdates/s",
        "cores": 1,
        "kind": kind,
        "sample": "CPU SIRT, linear projector, 256x256, 180 angles x 384 dets: median of %d runs "
                  "of %d iterations after a warm-up, %.2f s per run (%.1f ms/iteration; runs "
                  "%s s); updates = 2 x angles x dets x 256 per iteration"
                  % (runs, iters, dt, 1e3 * dt / iters, ", ".join("%.2f" % x for x in secs)),
        "seconds": dt,
        "host": host_cpu(),
    }


def _cpu_sirt_worker(job):
    """Independent CPU SIRT runs (one process of the all-cores baseline) on
    the phantom image `img`: a warm-up, then `runs` timed runs of `iters`
    iterations; returns their seconds.  Imports numpy and the CPU library
    only."""
    img, iters, runs = job
    from oracle import cpu

    lib = cpu.reference() or cpu.port()
    n, na, nd = 256, 180, 384
    ang = np.linspace(0, np.pi, na, endpoint=False)
    sino = lib.fp(img, ang, nd)
    lib.sirt(sino, ang, (n, n), 1)
    out = []
    for _ in range(runs):
        t0 = time.perf_counter()
        lib.sirt(sino, ang, (n, n), iters)
        out.append(time.perf_counter() - t0)
    return out


def cpu_all_cores(iters, workers, runs=5):
    """Throughput of `workers` concurrent independent CPU SIRT processes
    (the reference CPU path is single-threaded, so the host's cores are
    used as independent jobs), each timing `runs` runs of `iters`
    iterations; the rate uses the median over all runs.  Runs before the
    GPU is touched."""
    import multiprocessing as mp

    from astra import phantom

    img = phantom.ellipses_2d(256, seed=0)
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    # the workers finish and exit on their own (close + join): a pool torn
    # down with SIGTERM (the context manager's terminate) can leave workers
    # that a profiler's signal handler keeps alive, and the join then waits
    pool = ctx.Pool(workers)
    try:
        secs = pool.map_async(_cpu_sirt_worker, [(img, iters, runs)] * workers).get(timeout=900)
    finally:
        pool.close()
        pool.join()
    wall = time.perf_counter() - t0
    n, na, nd = 256, 180, 384
    per_run = float(np.median([x for w in secs for x in w]))
    updates = workers * iters * 2.0 * na * nd * n
    from oracle import cpu

    kind = "reference" if cpu.reference() is not None else "port"
    host = host_cpu()
    return {"value": updates / per_run / 1e9, "unit": "Gupdates/s", "cores": workers, "kind": kind,
            "host": host,
            "sample": "%d concurrent CPU SIRT processes (256x256, 180 x 384), each %d runs of %d "
                      "iterations after a warm-up; median run %.2f s, slowest %.2f s; pool wall "
                      "time incl. start-up %.2f s; host nproc %s, %s CPUs in this process's "
                      "affinity, %s"
                      % (workers, runs, iters, per_run, max(x for w in secs for x in w), wall,
                         host["nproc"], host["affinity_cpus"], host["model"])}


def copy_calibration(dev, nbytes=1 << 30):
    """One plain device-to-device copy of `nbytes` (a kernel of known HBM
    traffic: nbytes read + nbytes written) used to scale the profiler's
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize(dev)
    b.copy_(a)
    torch.cuda.synchronize(dev)
    log("bench: copy calibration of %d bytes done" % nbytes)
    del a, b
    torch.cuda.empty_cache()


def batched_sirt_bench(n_slices, n_plain, iters, host_phantoms=False, persist_iters=10):
    """Batched 2D SIRT slices/s (one hipGraph launch for all iterations of all
    slices) vs the plain per-slice SIRT_CUDA drop-in (one algorithm object
    per slice, run(iters), host data) on slices of the same seeded
    distribution: N log-uniform in [128, 1024], 90-720 angles, 1.5 N dets,
    parallel / fan alternating."""
    import torch

    import astra
    from astra import batched, phantom

    dev = torch.device("cuda", 0)
    slices = batched.config5_slices(n_slices, seed=0)
    pgs, vgs = batched.geometries(slices)
    B = batched.BatchedSIRT2D(pgs, vgs)
    truth = B.empty_volumes()
    if host_phantoms:
        # built on the host and uploaded once (a profiled run then sees a
        # few hundred kernel dispatches instead of tens of thousands)
        host = np.concatenate([phantom.ellipses_2d(s["n"], seed=1000 + i).astype(np.float32).ravel()
                               for i, s in enumerate(slices)])
        truth.copy_(torch.from_numpy(host))
    else:
        for i, s in enumerate(slices):
            B.volume(truth, i).copy_(phantom.ellipses_2d_device(s["n"], seed=1000 + i, device=dev))
    sinos = B.forward(truth)
    work = [iters * 2.5 * s["na"] * s["n"] * s["n"] for s in slices]  # ~FP 1.5 + BP 1 per pixel-angle
    x = B.empty_volumes()
    t = time.perf_counter()
    B.run(sinos, x, iters)  # autotune + weights + graph capture + first replay
    setup_s = time.perf_counter() - t
    x.zero_()
    torch.cuda.synchronize(dev)
    t = time.perf_counter()
    B.run(sinos, x, iters)  # one graph launch
    batched_s = time.perf_counter() - t
    log("bench: batched SIRT %d slices x %d iterations in %.3f s (first run incl. capture %.3f s)"
        % (n_slices, iters, batched_s, setup_s))
    # the same batch as ONE persistent launch per run (device work queues):
    # one warm-up iteration (queues + capture), then a timed run of
    # persist_iters iterations; the result must equal the graph's
    persist = None
    if persist_iters > 0:
        y = B.empty_volumes()
        B.run(sinos, y, 1, persistent=True)
        y.zero_()
        torch.cuda.synchronize(dev)
        t = time.perf_counter()
        B.run(sinos, y, persist_iters, persistent=True)
        persist_s = time.perf_counter() - t
        same = None
        if persist_iters == iters:
            same = bool(torch.equal(x, y))
        per_it = persist_s / persist_iters
        persist = {"seconds_per_iteration": per_it,
                   "slices_per_s": n_slices / (per_it * iters),
                   "graph_seconds_per_iteration": batched_s / iters,
                   "timed_iterations": persist_iters,
                   "equal_to_graph": same}
        log("bench: persistent batched SIRT %.1f ms per iteration (graph %.1f)"
            % (1e3 * per_it, 1e3 * batched_s / iters))
        del y

    # plain path on n_plain slices spread evenly over the same list (every
    # (n_slices / n_plain)-th: the same size and angle distribution)
    stride = max(1, n_slices // max(n_plain, 1))
    picks = list(range(0, n_slices, stride))[:n_plain]
    n_plain = len(picks)
    host_sinos = {i: B.sinogram(sinos, i).cpu().numpy() for i in picks}
    torch.cuda.synchronize(dev)
    t = time.perf_counter()
    for i in picks:
        pid = astra.create_projector("cuda", pgs[i], vgs[i])
        sid = astra.data2d.create("-sino", pgs[i], host_sinos[i])
        rid = astra.data2d.create("-vol", vgs[i], 0)
        cfg = astra.astra_dict("SIRT_CUDA")
        cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
        aid = astra.algorithm.create(cfg)
        astra.algorithm.run(aid, iters)
        astra.data2d.get(rid)
        astra.algorithm.delete(aid)
        astra.data2d.delete([sid, rid])
        astra.projector.delete(pid)
    plain_s = time.perf_counter() - t
    log("bench: plain per-slice SIRT_CUDA %d slices in %.3f s" % (n_plain, plain_s))
    b_rate = n_slices / batched_s
    p_rate = n_plain / plain_s
    b_upd = sum(work) / batched_s
    p_upd = sum(work[i] for i in picks) / plain_s
    del B, truth, sinos, x
    return {"slices_per_s": b_rate, "plain_slices_per_s": p_rate, "speedup": b_rate / p_rate,
            "work_normalised_speedup": b_upd / p_upd, "batched_Gupdates_per_s": b_upd / 1e9,
            "plain_Gupdates_per_s": p_upd / 1e9, "slices": n_slices, "plain_slices": n_plain,
            "plain_sample": "every %d-th slice of the batch" % stride,
            "iterations": iters, "batched_seconds": batched_s, "plain_seconds": plain_s,
            "first_run_seconds": setup_s, "persistent": persist}


def differentiable_bench(n=256, na=360, nd=384, iters=20):
    """Differentiable cone projector: one forward projection + matched-adjoint
    gradient of 1/2 ||A x - y||^2, and a 20-iteration unrolled SIRT with
    the non-negativity constraint (forward + backward w.r.t. the
    projections) with recomputed iterates vs autograd through stored
    iterates (time and peak torch memory)."""
    import torch

    import astra
    from astra import autograd as AG, phantom

    dev = torch.device("cuda", 0)
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, ang, 1000.0, 500.0)
    pid = astra.create_projector("cuda3d", pg, vg)
    truth = phantom.ellipsoids_3d_device(n, count=20, seed=0, device=dev)
    y = AG.forward_project(truth, pid).detach()
    target = truth.clone()

    def grad_step():
        x = torch.zeros_like(truth).requires_grad_(True)
        loss = 0.5 * ((AG.forward_project(x, pid) - y) ** 2).sum()
        loss.backward()
        return x.grad

    grad_step()
    torch.cuda.synchronize(dev)
    t = time.perf_counter()
    for _ in range(3):
        grad_step()
    torch.cuda.synchronize(dev)
    grad_s = (time.perf_counter() - t) / 3

    W = AG.sirt_weights(pid)
    res = {}
    from astra import _lib
    keep = lambda *a, **k: AG.sirt_unrolled(*a, keep_segment=True, **k)
    for name, fn in (("recompute", AG.sirt_unrolled), ("recompute_keep_segment", keep),
                     ("stored", AG.sirt_unrolled_stored)):
        b = y.clone().requires_grad_(True)
        torch.cuda.synchronize(dev)
        torch.cuda.reset_peak_memory_stats(dev)
        base = torch.cuda.memory_allocated(dev)
        sbase, _ = _lib.scratch_usage(reset_peak=True)
        t = time.perf_counter()
        x = fn(b, pid, iters, weights=W, min_constraint=0.0)
        loss = ((x - target) ** 2).sum()
        loss.backward()
        torch.cuda.synchronize(dev)
        secs = time.perf_counter() - t
        _, speak = _lib.scratch_usage()
        torch_extra = (torch.cuda.max_memory_allocated(dev) - base) / 2 ** 20
        scratch_extra = (speak - sbase) / 2 ** 20
        # the two peaks need not coincide in time: their sum bounds the device peak
        res[name] = {"seconds": secs, "peak_extra_MiB": torch_extra,
                     "library_scratch_peak_MiB": scratch_extra,
                     "device_peak_extra_MiB_bound": torch_extra + scratch_extra}
        del b, x, loss
    log("bench: differentiable %d^3: grad step %.3f s; unrolled SIRT %d its fwd+bwd recompute %.2f s "
        "(+%.0f MiB torch, +%.0f MiB scratch), recompute keeping segments %.2f s (+%.0f MiB torch) "
        "vs stored %.2f s (+%.0f MiB torch, +%.0f MiB scratch)"
        % (n, grad_s, iters, res["recompute"]["seconds"], res["recompute"]["peak_extra_MiB"],
           res["recompute"]["library_scratch_peak_MiB"], res["recompute_keep_segment"]["seconds"],
           res["recompute_keep_segment"]["peak_extra_MiB"], res["stored"]["seconds"],
           res["stored"]["peak_extra_MiB"], res["stored"]["library_scratch_peak_MiB"]))
    astra.projector3d.delete(pid)
    return {"workload": "cone %d^3, %d angles, %dx%d detector, SOD 1000, ODD 500" % (n, na, nd, nd),
            "fp_plus_matched_grad_seconds": grad_s, "unrolled_sirt_iterations": iters,
            "unrolled_sirt": res}


def fan2d_bench(sirt_iters=200):
    """2D fan beam at 1024^2: seeded 10-ellipse phantom, 720 angles over
    [0, 2pi), 1536 detectors of width 1, SOD 1000, ODD 500, Poisson noise
    for 1e5 incident photons; FBP (Ram-Lak) and SIRT_CUDA (200 iterations),
    data linked on the GPU (DLPack), seconds per job after one warm-up."""
    import torch

    import astra
    from astra import phantom

    n, na, nd = 1024, 720, 1536
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n)
    pg = astra.create_proj_geom("fanflat", 1.0, nd, ang, 1000.0, 500.0)
    pid = astra.create_projector("cuda", pg, vg)
    _, clean = astra.create_sino(phantom.ellipses_2d(n, seed=0), pid)
    noisy = astra.add_noise_to_sino(clean, 1e5, seed=0).astype(np.float32)
    sino_t = torch.from_numpy(noisy).cuda()
    rec_t = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    sid = astra.data2d.link("-sino", pg, sino_t)
    rid = astra.data2d.link("-vol", vg, rec_t)
    res = {"workload": "fan 1024^2, 720 angles over [0,2pi), 1536 dets, SOD 1000, ODD 500, "
                       "Poisson 1e5 photons"}
    for name, iters in (("FBP_CUDA", 1), ("SIRT_CUDA", sirt_iters)):
        cfg = astra.astra_dict(name)
        cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
        aid = astra.algorithm.create(cfg)
        astra.algorithm.run(aid, 1 if iters == 1 else 2)  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        astra.algorithm.run(aid, iters)
        torch.cuda.synchronize()
        res[name.split("_")[0].lower() + "_seconds"] = time.perf_counter() - t
        astra.algorithm.delete(aid)
    fp_samples = float(na) * nd * n
    bp_updates = float(n) * n * na
    res["sirt_iterations"] = sirt_iters
    res["sirt_Gupdates_per_s"] = sirt_iters * (fp_samples + bp_updates) / res["sirt_seconds"] / 1e9
    log("bench: fan 1024^2: FBP %.4f s; SIRT %d its %.3f s (%.0f Gupdates/s)"
        % (res["fbp_seconds"], sirt_iters, res["sirt_seconds"], res["sirt_Gupdates_per_s"]))
    astra.data2d.delete([sid, rid])
    astra.projector.delete(pid)
    return res


def sirt3d_bench(n=512, na=720, nd=768, iters=50):
    """SIRT3D_CUDA, 50 iterations on the 512^3 / 720 x 768^2 cone workload
    (projections = FP of the 20-ellipsoid phantom), data linked on the GPU,
    after a 2-iteration warm-up on the same algorithm object."""
    import torch

    import astra
    from astra import _core, _lib, phantom

    dev = torch.device("cuda", 0)
    angles = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, angles, 1000.0, 500.0)
    pid = astra.create_projector("cuda3d", pg, vg)
    P = astra.projector3d.get(pid)
    truth = phantom.ellipsoids_3d_device(n, count=20, seed=0, device=dev)
    proj = torch.empty((nd, na, nd), dtype=torch.float32, device=dev)
    _core.fp(P, truth, proj, _lib.MODE_SET)
    del truth
    rec = torch.zeros((n, n, n), dtype=torch.float32, device=dev)
    sid = astra.data3d.link("-sino", pg, proj)
    rid = astra.data3d.link("-vol", vg, rec)
    cfg = astra.astra_dict("SIRT3D_CUDA")
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid)
    aid = astra.algorithm.create(cfg)
    astra.algorithm.run(aid, 2)  # warm-up: weights, copies, scratch pool
    rec.zero_()
    torch.cuda.synchronize(dev)
    t = time.perf_counter()
    astra.algorithm.run(aid, iters)
    torch.cuda.synchronize(dev)
    dt = time.perf_counter() - t
    res_norm = astra.algorithm.get_res_norm(aid)
    astra.algorithm.delete(aid)
    astra.data3d.delete([sid, rid])
    astra.projector3d.delete(pid)
    del proj, rec
    torch.cuda.empty_cache()
    ups = iters * (float(na) * nd * nd * n + float(n) ** 3 * na) / dt / 1e9
    log("bench: SIRT3D %d^3, %d iterations in %.3f s (%.1f ms/iteration, %.0f Gupdates/s)"
        % (n, iters, dt, 1e3 * dt / iters, ups))
    return {"workload": "cone %d^3, %d angles, %dx%d detector, SOD 1000, ODD 500" % (n, na, nd, nd),
            "iterations": iters, "seconds": dt, "ms_per_iteration": 1e3 * dt / iters,
            "Gupdates_per_s": ups, "residual_norm": res_norm}


def large_cone_bench(n=1024, na=1440, nd=1536, steps=2):
    """Cone FP+BP and FDK at 1024^3 / 1440 x 1536^2, unchunked on one GPU
    (volume, projections and the internal copies resident in HBM)."""
    import torch

    import astra
    from astra import _core, _lib, phantom
    from astra.algorithm import fdk

    dev = torch.device("cuda", 0)
    angles = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, angles, 1000.0, 500.0)
    pid = astra.create_projector("cuda3d", pg, vg)
    P = astra.projector3d.get(pid)
    vol = phantom.ellipsoids_3d_device(n, count=20, seed=0, device=dev)
    proj = torch.empty((nd, na, nd), dtype=torch.float32, device=dev)
    bpv = torch.empty_like(vol)

    def step():
        _core.fp(P, vol, proj, _lib.MODE_SET)
        _core.bp(P, proj, bpv, _lib.MODE_SET)

    step()
    torch.cuda.synchronize(dev)
    t = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize(dev)
    dt = (time.perf_counter() - t) / steps
    pcopy = proj.clone()
    fdk(P, pcopy, bpv, _lib.MODE_SET)  # warm-up (weights, filter, scratch pool)
    pcopy.copy_(proj)
    torch.cuda.synchronize(dev)
    t = time.perf_counter()
    fdk(P, pcopy, bpv, _lib.MODE_SET)
    torch.cuda.synchronize(dev)
    fdk_s = time.perf_counter() - t
    pcopy.copy_(proj)
    _lib.profile_enable(True)
    fdk(P, pcopy, bpv, _lib.MODE_SET)
    fms, _, fl = _lib.profile_read("filter_rows")
    _lib.profile_enable(False)
    peak = torch.cuda.max_memory_allocated(dev) / 2 ** 30
    ups = (float(na) * nd * nd * n + float(n) ** 3 * na) / dt / 1e9
    log("bench: cone 1024^3 / 1440 x 1536^2: FP+BP %.3f s/step (%.0f Gupdates/s), FDK %.3f s "
        "(row filter %.1f ms)" % (dt, ups, fdk_s, fms / max(fl, 1)))
    del vol, proj, bpv, pcopy
    astra.projector3d.delete(pid)
    torch.cuda.empty_cache()
    return {"workload": "cone %d^3, %d angles, %dx%d detector, SOD 1000, ODD 500" % (n, na, nd, nd),
            "fpbp_seconds_per_step": dt, "Gupdates_per_s": ups, "fdk_seconds": fdk_s,
            "fdk_row_filter_ms": fms / max(fl, 1),
            "torch_peak_GiB": peak}


TRAFFIC_FILE = "profiles/cone512_fpbp_counters.md"


def profile_value(key, kernel, n, na, nd):
    """A per-launch counter figure of `kernel` at the 512^3 / 720 x 768^2
    workload recorded in TRAFFIC_FILE as a line `<key> <kernel> <value>`:
    `traffic_bytes_per_launch` (calibrated FETCH_SIZE + WRITE_SIZE) or
    `vmem_rd_insts_per_launch` (SQ_INSTS_VMEM_RD).  Profiler counters
    cannot be read from inside this run; None for any other workload."""
    if (n, na, nd) != (512, 720, 768):
        return None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), TRAFFIC_FILE)
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == key and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        return None
    return None


def gather_limits():
    """Lane-loads per second of the 16-byte gather calibration kernel (the
    taps' lane pattern, and coalesced), measured on this device now."""
    from astra import _lib

    return {"taps": _lib.calibrate_gather(0), "coalesced": _lib.calibrate_gather(1),
            "taps_2d_8byte": _lib.calibrate_gather(2),
            "tap_pairs_2d_16byte_align4": _lib.calibrate_gather(3),
            "tap_pairs_2d_16byte_align8": _lib.calibrate_gather(4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--angles", type=int, default=720)
    ap.add_argument("--det", type=int, default=768)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fdk", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=50,
                    help="SIRT iterations per timed CPU run (median of 5 runs)")
    ap.add_argument("--batched-slices", type=int, default=4096)
    ap.add_argument("--plain-slices", type=int, default=128)
    ap.add_argument("--batched-iters", type=int, default=50)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--persist-iters", type=int, default=10,
                    help="timed iterations of the persistent batched run (0: skip)")
    ap.add_argument("--batched-host-phantoms", action="store_true",
                    help="build the batched workload's phantoms on the host (for profiled runs)")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--diff-iters", type=int, default=20,
                    help="unrolled SIRT iterations of the differentiable leg (fewer for profiled runs)")
    ap.add_argument("--no-fan2d", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-sirt3d", action="store_true")
    ap.add_argument("--sirt3d-iters", type=int, default=50)
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="concurrent CPU SIRT processes for the all-cores figure (0: the CPUs "
                         "this process may use, at most 16: a one-GPU box's CPU share)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device copy first (profiler counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the projection path is measured on a single "
              "MI355X (multi-GPU sharding is design-only); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    cpu_all = None
    if not args.no_cpu:
        workers = args.cpu_workers or min(16, host_cpu()["affinity_cpus"])
        log("bench: CPU baseline, %d concurrent processes" % workers)
        cpu_all = cpu_all_cores(max(args.cpu_iters // 2, 1), workers)

    import torch

    import astra
    from astra import _core, _lib, phantom
    from astra.algorithm import fdk

    dev = torch.device("cuda", 0)
    if args.copy_calibration:
        copy_calibration(dev)
    n, na, nd = args.size, args.angles, args.det
    sod, odd = 1000.0, 500.0
    angles = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vg = astra.create_vol_geom(n, n, n)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, nd, nd, angles, sod, odd)
    pid = astra.create_projector("cuda3d", pg, vg)
    P = astra.projector3d.get(pid)

    log("bench: building %d^3 phantom on %s" % (n, torch.cuda.get_device_name(dev)))
    vol = phantom.ellipsoids_3d_device(n, count=20, seed=0, device=dev)
    proj = torch.empty((nd, na, nd), dtype=torch.float32, device=dev)
    bpv = torch.empty((n, n, n), dtype=torch.float32, device=dev)

    def step():
        _core.fp(P, vol, proj, _lib.MODE_SET)
        _core.bp(P, proj, bpv, _lib.MODE_SET)

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize(dev)
        log("bench: warmup step %d done" % i)
    _lib.profile_enable(True)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    torch.cuda.synchronize(dev)
    t1 = time.perf_counter()
    log("bench: %d timed steps in %.3f s" % (args.steps, t1 - t0))
    kernels = {}
    for name in ("fp3d_cone", "bp3d_cone"):
        ms, units, launches = _lib.profile_read(name)
        kernels[name] = {"total_ms": ms, "units": units, "launches": launches,
                         "avg_launch_ms": ms / max(launches, 1)}
    _lib.profile_enable(False)

    fp_samples = float(na) * nd * nd * n
    bp_updates = float(n) ** 3 * na
    elapsed = t1 - t0
    value = args.steps * (fp_samples + bp_updates) / elapsed / 1e9

    # roofline of the dominant kernel.  Its algorithmic bytes (16 B per
    # bilinear sample, 8 B output read-modify-write per ray / voxel) are
    # served mostly by L1 / L2, so they are reported beside the HBM figure,
    # not as it: `achieved` is the measured HBM traffic per launch
    # (calibrated profiler counters) over the live launch time.
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    for k, v in kernels.items():
        u = v["units"] / max(v["launches"], 1)
        o = u / (n if k.startswith("fp") else na)
        t = v["avg_launch_ms"] * 1e-3
        v["algorithmic_bytes_per_launch"] = 16.0 * u + 8.0 * o
        v["algorithmic_GBps"] = v["algorithmic_bytes_per_launch"] / t / 1e9
        v["Gupdates_per_s"] = u / t / 1e9
        tr = profile_value("traffic_bytes_per_launch", k, n, na, nd)
        v["hbm_traffic_bytes_per_launch"] = tr
        v["hbm_GBps"] = tr / t / 1e9 if tr else None
        v["hbm_frac"] = tr / t / 1e9 / HBM_PEAK_GBS if tr else None
        vm = profile_value("vmem_rd_insts_per_launch", k, n, na, nd)
        v["issued_lane_loads_per_launch"] = 64.0 * vm if vm else None
        v["issued_lane_loads_per_s"] = 64.0 * vm / t if vm else None
    kd = kernels[dom]
    limits = gather_limits()
    log("bench: gather calibration %.3e lane-loads/s (taps), %.3e (coalesced)"
        % (limits["taps"], limits["coalesced"]))
    issued = kd["issued_lane_loads_per_s"]
    if kd["hbm_GBps"] is not None:
        achieved, achieved_kind = kd["hbm_GBps"], "measured HBM traffic"
    else:
        achieved, achieved_kind = kd["algorithmic_GBps"], "algorithmic bytes (no counters for this workload)"
    roofline = {
        "bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": kd["hbm_traffic_bytes_per_launch"],
        "achieved_kind": achieved_kind,
        "traffic_source": TRAFFIC_FILE if kd["hbm_traffic_bytes_per_launch"] else None,
        "avg_launch_ms": kd["avg_launch_ms"],
        "algorithmic": {"bytes_per_launch": kd["algorithmic_bytes_per_launch"],
                        "GBps": kd["algorithmic_GBps"],
                        "note": "16 B per bilinear sample + 8 B per output; served by L1/L2, "
                                "so above the HBM peak"},
        "per_kernel_hbm_frac": {k: v["hbm_frac"] for k, v in kernels.items()},
        "texture_path": {
            "issued_lane_loads_per_launch": kd["issued_lane_loads_per_launch"],
            "issued_lane_loads_per_s": issued,
            "limit_lane_loads_per_s": limits["taps"],
            "coalesced_lane_loads_per_s": limits["coalesced"],
            "taps_2d_8byte_lane_loads_per_s": limits["taps_2d_8byte"],
            "tap_pairs_2d_16byte_align4_lane_loads_per_s": limits["tap_pairs_2d_16byte_align4"],
            "tap_pairs_2d_16byte_align8_lane_loads_per_s": limits["tap_pairs_2d_16byte_align8"],
            "ta_busy_frac": profile_value("ta_busy_frac", dom, n, na, nd),
            "frac": issued / limits["taps"] if issued else None,
            "note": "issued = SQ_INSTS_VMEM_RD x 64 per launch (profile) over the live launch "
                    "time; limit = astra_amd_calibrate_gather(mode 0) measured in this run: "
                    "16-byte loads of a 16 x 4 ray bundle two thirds of a texel apart, from "
                    "L1-resident data; ta_busy_frac = TA_BUSY_avr over the kernel's cycles "
                    "(profile)"},
    }

    fdk_s = None
    if not args.no_fdk:
        pcopy = proj.clone()
        out = torch.empty_like(vol)
        fdk(P, pcopy, out, _lib.MODE_SET)  # warm-up (filter + weights built once)
        pcopy.copy_(proj)
        torch.cuda.synchronize(dev)
        t = time.perf_counter()
        fdk(P, pcopy, out, _lib.MODE_SET)
        torch.cuda.synchronize(dev)
        fdk_s = time.perf_counter() - t
        # the row filter's own time, from a second (event-timed) run
        pcopy.copy_(proj)
        _lib.profile_enable(True)
        fdk(P, pcopy, out, _lib.MODE_SET)
        ms, rows, launches = _lib.profile_read("filter_rows")
        _lib.profile_enable(False)
        n_fft = 1
        while n_fft < 2 * nd:
            n_fft *= 2
        kernels["filter_rows"] = {
            "total_ms": ms, "units": rows, "launches": launches,
            "avg_launch_ms": ms / max(launches, 1),
            # rows read and written once (fp32), pre-weights read
            "hbm_bytes_per_launch": rows * nd * 8.0 / max(launches, 1),
            "GBps": rows * nd * 8.0 / max(ms * 1e-3, 1e-12) / 1e9,
            # 5 N log2 N per complex transform, two per row pair
            "GFLOPs": rows / 2.0 * 2 * 5.0 * n_fft * np.log2(n_fft) / max(ms * 1e-3, 1e-12) / 1e9}
        log("bench: FDK %.3f s (row filter %.2f ms)" % (fdk_s, ms / max(launches, 1)))
        del pcopy, out

    del proj, bpv
    torch.cuda.empty_cache()
    batched_res = None
    if not args.no_batched:
        batched_res = batched_sirt_bench(args.batched_slices, args.plain_slices, args.batched_iters,
                                         args.batched_host_phantoms, args.persist_iters)

    diff_res = None
    if not args.no_diff:
        diff_res = differentiable_bench(iters=args.diff_iters)
    fan_res = None if args.no_fan2d else fan2d_bench()
    sirt3d_res = None if args.no_sirt3d else sirt3d_bench(n, na, nd, args.sirt3d_iters)
    large_res = None if args.no_large else large_cone_bench()

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (%d SIRT iterations)" % args.cpu_iters)
        cpu = cpu_baseline(args.cpu_iters)
        cpu["all_cores"] = cpu_all

    result = {
        "metric": "cone FP+BP Gupdates/s & FDK sec at 512^3; batched 2D SIRT slices/s vs plain",
        "value": value,
        "unit": "Gupdates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (20 seeded ellipsoids, densities in [0,1])",
        "config": {"workload": "cone FP+BP, %d^3 volume, %d projections over [0,2pi), %dx%d "
                               "detector, pixel 1.0, SOD %g, ODD %g" % (n, na, nd, nd, sod, odd),
                   "volume": [n, n, n], "angles": na, "detector": [nd, nd],
                   "source_origin": sod, "origin_detector": odd},
        "roofline": roofline,
        "fdk_seconds": fdk_s,
        "batched_sirt": batched_res,
        "differentiable": diff_res,
        "fan2d_1024": fan_res,
        "sirt3d_512": sirt3d_res,
        "cone_1024": large_res,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
