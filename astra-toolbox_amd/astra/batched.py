"""Batched 2D SIRT: many independent 2D reconstructions in one GPU pipeline.

Each slice has its own 2D volume geometry and parallel / fan projection
geometry (sizes, angle counts and detector counts may all differ).  The
whole batch runs as one hipGraph launch per ``run()`` (all iterations
captured), with per-batch autotuned tile shapes; results equal running
``SIRT_CUDA`` on every slice separately.
"""
import ctypes

import numpy as np

from . import _core, _lib
from . import geometry as G
from .log import AstraError

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


class BatchedSIRT2D:
    def __init__(self, proj_geoms, vol_geoms, gpu_index=None):
        if len(proj_geoms) != len(vol_geoms) or not proj_geoms:
            raise AstraError("BatchedSIRT2D needs one volume geometry per projection geometry")
        self.n = len(proj_geoms)
        kinds, ws, hs, nas, nds, oscs, vecs = [], [], [], [], [], [], []
        self.vol_shapes, self.sino_shapes = [], []
        for pg, vg in zip(proj_geoms, vol_geoms):
            kind, v = G.vectors_2d(pg)
            v, osc = G.normalise_2d(kind, v, vg)
            h, w = G.vol_shape(vg)
            na, nd = G.proj_shape(pg)
            kinds.append(_lib.GEOM_FAN if kind == "fan" else _lib.GEOM_PAR)
            ws.append(w); hs.append(h); nas.append(na); nds.append(nd); oscs.append(osc)
            vecs.append(np.asarray(v, dtype=np.float32))
            self.vol_shapes.append((h, w))
            self.sino_shapes.append((na, nd))
        self.device = _core.device(gpu_index)
        u = lambda a: np.ascontiguousarray(a, dtype=np.uint32)
        self._arrays = dict(kinds=np.ascontiguousarray(kinds, dtype=np.int32), ws=u(ws), hs=u(hs),
                            nas=u(nas), nds=u(nds), vecs=np.ascontiguousarray(np.concatenate(vecs)),
                            oscs=np.ascontiguousarray(oscs, dtype=np.float32))
        a = self._arrays
        h = ctypes.c_void_p()
        p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        with torch.cuda.device(self.device):
            _lib.call("astra_amd_batch2d_create", self.n, p(a["kinds"]), p(a["ws"]), p(a["hs"]),
                      p(a["nas"]), p(a["nds"]), p(a["vecs"]), p(a["oscs"]), ctypes.byref(h))
        self._h = h
        self.vol_offsets = np.zeros(self.n + 1, dtype=np.uint64)
        self.sino_offsets = np.zeros(self.n + 1, dtype=np.uint64)
        _lib.call("astra_amd_batch2d_offsets", h, p(self.vol_offsets), p(self.sino_offsets))

    # ---------------------------------------------------------- packing --
    def empty_volumes(self):
        return torch.zeros(int(self.vol_offsets[-1]), dtype=torch.float32, device=self.device)

    def empty_sinograms(self):
        return torch.zeros(int(self.sino_offsets[-1]), dtype=torch.float32, device=self.device)

    def pack_volumes(self, vols):
        out = self.empty_volumes()
        for i, v in enumerate(vols):
            t = torch.as_tensor(v, dtype=torch.float32).reshape(-1)
            out[int(self.vol_offsets[i]):int(self.vol_offsets[i + 1])] = t.to(self.device)
        return out

    def pack_sinograms(self, sinos):
        out = self.empty_sinograms()
        for i, s in enumerate(sinos):
            t = torch.as_tensor(s, dtype=torch.float32).reshape(-1)
            out[int(self.sino_offsets[i]):int(self.sino_offsets[i + 1])] = t.to(self.device)
        return out

    def volume(self, packed, i):
        return packed[int(self.vol_offsets[i]):int(self.vol_offsets[i + 1])].view(self.vol_shapes[i])

    def sinogram(self, packed, i):
        return packed[int(self.sino_offsets[i]):int(self.sino_offsets[i + 1])].view(self.sino_shapes[i])

    # ---------------------------------------------------------- compute --
    def _check(self, t, n, what):
        if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != n or t.device != self.device:
            raise AstraError("%s must be a contiguous float32 tensor of %d elements on %s"
                             % (what, n, self.device))

    def forward(self, vols, sinos=None):
        """sinos = A vols for every slice (packed tensors)."""
        self._check(vols, int(self.vol_offsets[-1]), "volumes")
        if sinos is None:
            sinos = self.empty_sinograms()
        self._check(sinos, int(self.sino_offsets[-1]), "sinograms")
        with torch.cuda.device(self.device):
            _lib.call("astra_amd_batch2d_fp", self._h, _core.ptr(vols), _core.ptr(sinos),
                      _core.stream_ptr(self.device))
        return sinos

    def run(self, sinos, vols, iterations, relaxation=1.0, min_constraint=None,
            max_constraint=None, use_graph=True, sync=True, persistent=False):
        """`iterations` SIRT iterations on every slice; vols is updated in place.

        use_graph: the per-iteration launches (slice copies, FP, BP) are
        captured once into a hipGraph and replayed as one graph launch.
        persistent: the whole run is ONE persistent kernel launch over device
        work queues (it returns when the run has ended).  Every mode gives
        the same floats."""
        self._check(sinos, int(self.sino_offsets[-1]), "sinograms")
        self._check(vols, int(self.vol_offsets[-1]), "volumes")
        lo = -np.inf if min_constraint is None else float(min_constraint)
        hi = np.inf if max_constraint is None else float(max_constraint)
        with torch.cuda.device(self.device):
            _lib.call("astra_amd_batch2d_sirt", self._h, _core.ptr(sinos), _core.ptr(vols),
                      int(iterations), float(relaxation), lo, hi,
                      2 if persistent else (1 if use_graph else 0),
                      _core.stream_ptr(self.device))
            if sync:
                _core.sync(self.device)
        return vols

    def release(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            try:
                with torch.cuda.device(self.device):
                    _lib.call("astra_amd_batch2d_destroy", self._h)
            finally:
                self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def config5_slices(count, seed=0, n_min=128, n_max=1024, a_min=90, a_max=720):
    """Seeded slice list of the batched workload: N log-uniform in
    [n_min, n_max], angles uniform in [a_min, a_max], detectors 1.5 N,
    parallel / fan alternating by slice parity (fan: source-origin 2N,
    origin-detector N)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        na = int(rng.integers(a_min, a_max + 1))
        nd = int(round(1.5 * n))
        out.append(dict(n=n, na=na, nd=nd, kind="parallel" if i % 2 == 0 else "fanflat"))
    return out


def geometries(slices):
    """(proj_geoms, vol_geoms) for a config5_slices list."""
    from .creators import create_proj_geom, create_vol_geom
    pgs, vgs = [], []
    for s in slices:
        n, na, nd = s["n"], s["na"], s["nd"]
        vgs.append(create_vol_geom(n, n))
        if s["kind"] == "parallel":
            pgs.append(create_proj_geom("parallel", 1.0, nd, np.linspace(0, np.pi, na, endpoint=False)))
        else:
            pgs.append(create_proj_geom("fanflat", 1.0, nd, np.linspace(0, 2 * np.pi, na, endpoint=False),
                                        2.0 * n, 1.0 * n))
    return pgs, vgs
