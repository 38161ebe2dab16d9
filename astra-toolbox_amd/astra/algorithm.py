"""Algorithm objects (python/astra/algorithm.py:28-90).

Types built here (include/astra/AlgorithmTypelist.h:62-90): FP_CUDA,
BP_CUDA, FBP_CUDA, SIRT_CUDA, CGLS_CUDA, FP3D_CUDA, BP3D_CUDA, FDK_CUDA,
SIRT3D_CUDA, CGLS3D_CUDA.  Each mirrors the configuration keys of the
reference algorithm class and runs through libastra_amd on one GPU.
"""
import contextlib
import ctypes
import math
import signal
import threading

import numpy as np

from . import _core, _lib, composite
from . import geometry as G
from ._registry import algorithms as _reg
from ._registry import data2d as _d2, data3d as _d3
from ._registry import projectors2d as _p2, projectors3d as _p3
from .log import AstraError

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


# Ctrl-C during an iterative run.  The reference checks shouldAbort() before
# every iteration (src/CudaSirtAlgorithm.cpp:347) and its MATLAB interface
# points the hook at the interpreter's interrupt flag
# (matlab/mex/astra_mex_algorithm_c.cpp:133).  Here a SIGINT handler,
# installed for the run's duration on the main thread, sets a flag that the
# native abort hook reads before each iteration; the run then stops, the
# iterate reached is written back and KeyboardInterrupt is raised.
_stop = threading.Event()


def _should_abort():
    return _stop.is_set()


def _hook(user):
    try:
        return 1 if _should_abort() else 0
    except BaseException:
        return 1


_native_hook = _lib.ABORT_HOOK(_hook)
_no_hook = _lib.ABORT_HOOK()  # NULL


@contextlib.contextmanager
def _interruptible():
    """Let SIGINT stop a native iterative run between iterations (main
    thread only; an ignored SIGINT or a handler installed outside Python is
    left alone)."""
    if threading.current_thread() is not threading.main_thread() or \
            signal.getsignal(signal.SIGINT) in (None, signal.SIG_IGN):
        yield
        return
    _stop.clear()
    prev = signal.signal(signal.SIGINT, lambda signum, frame: _stop.set())
    try:
        _lib.call("astra_amd_set_abort_hook", _native_hook, None)
        yield
    finally:
        _lib.call("astra_amd_set_abort_hook", _no_hook, None)
        signal.signal(signal.SIGINT, prev)


def _opt(cfg, key, default=None):
    for sect in ("option", "options", "Option", "Options"):
        o = cfg.get(sect) or {}
        if key in o:
            return o[key]
    return cfg.get(key, default)


def _truthy(v):
    if isinstance(v, str):
        return v.lower() in ("yes", "true", "1", "on")
    return bool(v)


class _Algo:
    is3d = False

    def __init__(self, cfg):
        self.cfg = cfg
        reg = _d3 if self.is3d else _d2
        self.sino = reg.get(cfg["ProjectionDataId"]) if "ProjectionDataId" in cfg else None
        vid = cfg.get("ReconstructionDataId", cfg.get("VolumeDataId"))
        self.vol = reg.get(vid) if vid is not None else None
        if self.sino is None or self.vol is None:
            raise AstraError("algorithm needs ProjectionDataId and a volume data id")
        opts = {}
        gi = _opt(cfg, "GPUindex", _opt(cfg, "GPUIndex"))
        if gi is not None:
            opts["GPUindex"] = gi
        for k in ("DetectorSuperSampling", "VoxelSuperSampling", "PixelSuperSampling"):
            v = _opt(cfg, k)
            if v is not None:
                opts["VoxelSuperSampling" if k == "PixelSuperSampling" else k] = v
        pid = cfg.get("ProjectorId")
        if pid is not None:
            self.proj = (_p3 if self.is3d else _p2).get(pid)
            if opts.get("GPUindex") is not None:
                self.proj = _core.Projector(self.proj.type, self.proj.proj_geom, self.proj.vol_geom,
                                            dict(self.proj.options, **opts))
        else:
            self.proj = _core.projector_for(self.sino.geom, self.vol.geom, opts)
        if tuple(self.proj.proj_shape) != self.sino.shape or tuple(self.proj.vol_shape) != self.vol.shape:
            raise AstraError("data objects do not match the projector geometry")

    def run(self, iterations):
        raise NotImplementedError

    def get_res_norm(self):
        raise AstraError("this algorithm has no residual norm")

    def description(self):
        return type(self).__name__

    def release(self):
        pass


# ---------------------------------------------------------- FP / BP -------
class _FP(_Algo):
    def run(self, iterations=1):
        if self.is3d:  # CCompositeGeometryManager job: split when it does not fit
            composite.fp(self.proj, self.vol, self.sino, _lib.MODE_SET)
            return
        dev = self.proj.device()
        v = self.vol.on_device(dev)
        s = self.sino.on_device(dev, init=False)
        _core.fp(self.proj, v, s, _lib.MODE_SET)
        self.sino.write_back(s)
        _core.sync(dev)


class _BP(_Algo):
    def run(self, iterations=1):
        if self.is3d:
            composite.bp(self.proj, self.sino, self.vol, _lib.MODE_SET)
            return
        dev = self.proj.device()
        s = self.sino.on_device(dev)
        v = self.vol.on_device(dev, init=False)
        _core.bp(self.proj, s, v, _lib.MODE_SET)
        self.vol.write_back(v)
        _core.sync(dev)


class FP_CUDA(_FP):
    pass


class BP_CUDA(_BP):
    pass


class FP3D_CUDA(_FP):
    is3d = True


class BP3D_CUDA(_BP):
    is3d = True


# ------------------------------------------------------------- FBP / FDK --
# filter names of convertStringToFilter (src/Filters.cpp:509-545) -> window
_WINDOWS = {"ram-lak": 0, "ramlak": 0, "shepp-logan": 1, "cosine": 2, "hamming": 3, "hann": 4,
            "tukey": 5, "lanczos": 6, "triangular": 7, "gaussian": 8, "barlett-hann": 9,
            "blackman": 10, "nuttall": 11, "blackman-harris": 12, "blackman-nuttall": 13,
            "flat-top": 14, "kaiser": 15, "parzen": 16}
# user filters: name -> (real space, one row per angle)
_CUSTOM = {"projection": (0, False), "sinogram": (0, True), "rprojection": (1, False),
           "rsinogram": (1, True)}


def filter_config(cfg):
    """FilterType / FilterD / FilterParameter / FilterSinogramId, read from
    cfg['option'] or cfg itself (getFilterConfigForAlgorithm,
    src/Filters.cpp:548-640).  Returns (type, d, parameter, custom array)."""
    ftype = str(_opt(cfg, "FilterType", "ram-lak")).lower()
    if ftype not in _WINDOWS and ftype not in _CUSTOM and ftype != "none":
        raise AstraError("Failed to convert \"%s\" into a filter." % ftype)
    d = float(_opt(cfg, "FilterD", 1.0))
    param = float(_opt(cfg, "FilterParameter", -1.0))
    data = None
    if ftype in _CUSTOM:
        fid = _opt(cfg, "FilterSinogramId", None)
        if fid is None:
            raise AstraError("FilterType '%s' needs FilterSinogramId" % ftype)
        from . import data2d
        data = np.ascontiguousarray(data2d.get(int(fid)), dtype=np.float32)
        if data.ndim == 1:
            data = data.reshape(1, -1)
    return ftype, d, param, data


def _filter_handle(n_det, n_angles, fc):
    """Native filter handle for a filter_config tuple (the caller destroys
    it); None means the built-in Ram-Lak (d = 1)."""
    ftype, d, param, data = fc
    if ftype in ("ram-lak", "ramlak") and d == 1.0:
        return None
    h = ctypes.c_void_p()
    if ftype == "none":
        _lib.call("astra_amd_filter_create", n_det, None, 0, ctypes.byref(h))
        return h
    if ftype in _CUSTOM:
        real, per_angle = _CUSTOM[ftype]
        rows = data.shape[0]
        if rows != (n_angles if per_angle else 1):
            raise AstraError("FilterSinogramId for '%s' must have %d row(s), got %d"
                             % (ftype, n_angles if per_angle else 1, rows))
        _lib.call("astra_amd_filter_create_custom", n_det, rows, real,
                  data.ctypes.data_as(ctypes.c_void_p), data.shape[1], ctypes.byref(h))
        return h
    n_fft = 1
    while n_fft < 2 * n_det:
        n_fft *= 2
    nb = n_fft // 2 + 1
    spec = np.zeros(nb, dtype=np.float32)
    _lib.call("astra_amd_filter_spectrum", _WINDOWS[ftype], n_det, d, param,
              spec.ctypes.data_as(ctypes.c_void_p), nb)
    _lib.call("astra_amd_filter_create", n_det, spec.ctypes.data_as(ctypes.c_void_p), nb,
              ctypes.byref(h))
    return h


class FBP_CUDA(_Algo):
    """CCudaFilteredBackProjectionAlgorithm (src/CudaFilteredBackProjectionAlgorithm.cpp:174-239)."""

    def __init__(self, cfg):
        super().__init__(cfg)
        self.short_scan = 1 if _truthy(_opt(cfg, "ShortScan", False)) else 0
        self.filter = filter_config(cfg)
        self._filters = {}

    def run(self, iterations=1):
        dev = self.proj.device()
        px, py = G.pixel_lengths(self.vol.geom)
        scale = float(self.proj.out_scale) / float(px * py)
        s = self.sino.on_device(dev).clone()
        v = self.vol.on_device(dev, init=False)
        h = ctypes.c_void_p(self.proj.handle(dev))
        with torch.cuda.device(dev):
            f = _cached_filter(self._filters, dev, self.proj.proj_shape[1], self.proj.proj_shape[0],
                               self.filter)
            _lib.call("astra_amd_fbp2d", h, _core.ptr(s), _core.ptr(v), scale, self.short_scan,
                      f, _lib.MODE_SET, _core.stream_ptr(dev))
        self.vol.write_back(v)
        _core.sync(dev)

    def release(self):
        _release_filters(self._filters)


class FDK_CUDA(_Algo):
    """CCudaFDKAlgorithm3D (src/CudaFDKAlgorithm3D.cpp:183-198) via one
    FDK job (split over volume slabs when it does not fit)."""
    is3d = True

    def __init__(self, cfg):
        super().__init__(cfg)
        self.short_scan = _truthy(_opt(cfg, "ShortScan", False))
        self.filter = filter_config(cfg)
        self._filters = {}
        if self.proj.kind != "cone":
            raise AstraError("FDK_CUDA requires a cone geometry")

    def run(self, iterations=1, mode=_lib.MODE_SET):
        nv, na, nu = self.proj.proj_shape
        dev = self.proj.device()
        with torch.cuda.device(dev):
            f = _cached_filter(self._filters, dev, nu, na, self.filter)
        composite.fdk(self.proj, self.sino, self.vol, mode, f, self.short_scan)

    def release(self):
        _release_filters(self._filters)


def _cached_filter(cache, dev, n_det, n_angles, fc):
    """The algorithm's native filter on `dev`, built on first use (None:
    the library's default Ram-Lak)."""
    if fc is None:
        return None
    if dev.index not in cache:
        cache[dev.index] = _filter_handle(n_det, n_angles, fc)
    return cache[dev.index]


def _release_filters(cache):
    for dev, f in list(cache.items()):
        if f is not None:
            try:
                with torch.cuda.device(dev):
                    _core.sync(torch.device("cuda", dev))
                    _lib.call("astra_amd_filter_destroy", f)
            except Exception:
                pass
    cache.clear()


def fdk(proj, p, v, mode, filter=None, short_scan=False):
    """FDK of device tensor p (overwritten by the filtered data) into v;
    `filter` is a filter_config tuple (None: Ram-Lak); short_scan adds
    Parker weights."""
    dev = v.device
    h = ctypes.c_void_p(proj.handle(dev))
    nv, na, nu = proj.proj_shape
    with torch.cuda.device(dev):
        f = None if filter is None else _filter_handle(nu, na, filter)
        try:
            _lib.call("astra_amd_fdk", h, _core.ptr(p), _core.ptr(v), 1.0, 1 if short_scan else 0,
                      f, mode, _core.stream_ptr(dev))
            _core.sync(dev)
        finally:
            if f is not None:
                _lib.call("astra_amd_filter_destroy", f)


# ------------------------------------------------------- SIRT / CGLS -----
class _Iterative(_Algo):
    algo_type = _lib.ALGO_SIRT

    def __init__(self, cfg):
        super().__init__(cfg)
        reg = _d3 if self.is3d else _d2
        self.masks = {}
        for key, name in (("ReconstructionMaskId", "ReconstructionMask"),
                          ("SinogramMaskId", "SinogramMask"),
                          ("MinMaskId", "MinMask"), ("MaxMaskId", "MaxMask")):
            mid = _opt(cfg, key)
            if mid is not None:
                self.masks[name] = reg.get(mid)
        self.options = {}
        if self.algo_type == _lib.ALGO_SIRT:
            self.options["Relaxation"] = float(_opt(cfg, "Relaxation", 1.0))
        # MinConstraint / MaxConstraint, or the deprecated Use*/..Value pair
        # (src/ReconstructionAlgorithm2D.cpp:105-127)
        if _opt(cfg, "MinConstraint") is not None:
            self.options["MinConstraint"] = float(_opt(cfg, "MinConstraint"))
        elif _truthy(_opt(cfg, "UseMinConstraint", False)):
            self.options["MinConstraint"] = float(_opt(cfg, "MinConstraintValue", 0.0))
        if _opt(cfg, "MaxConstraint") is not None:
            self.options["MaxConstraint"] = float(_opt(cfg, "MaxConstraint"))
        elif _truthy(_opt(cfg, "UseMaxConstraint", False)):
            self.options["MaxConstraint"] = float(_opt(cfg, "MaxConstraintValue", 255.0))
        self._handles = {}
        self._keep = {}

    def _handle(self, dev):
        h = self._handles.get(dev.index)
        if h is not None:
            return h
        g = ctypes.c_void_p(self.proj.handle(dev))
        out = ctypes.c_void_p()
        with torch.cuda.device(dev):
            if self.is3d:
                p = self.proj.params3d()
                _lib.call("astra_amd_algo3d_create", self.algo_type, g, ctypes.byref(p), ctypes.byref(out))
            else:
                p = self.proj.params2d()
                _lib.call("astra_amd_algo2d_create", self.algo_type, g, ctypes.byref(p), ctypes.byref(out))
        for k, v in self.options.items():
            _lib.call("astra_amd_algo_set_option", out, k.encode(), float(v))
        self._handles[dev.index] = out
        return out

    def _bind_masks(self, h, dev):
        keep = {}
        for name, d in self.masks.items():
            t = d.on_device(dev)
            keep[name] = t
            _lib.call("astra_amd_algo_set_mask", h, name.encode(), _core.ptr(t))
        self._keep[dev.index] = keep

    def run(self, iterations=1):
        dev = self.proj.device()
        h = self._handle(dev)
        with torch.cuda.device(dev):
            self._bind_masks(h, dev)
            b = self.sino.on_device(dev)
            x = self.vol.on_device(dev)
            try:
                with _interruptible():
                    _lib.call("astra_amd_algo_run", h, _core.ptr(b), _core.ptr(x),
                              int(iterations), _core.stream_ptr(dev))
            except _lib.Aborted as e:
                self.vol.write_back(x)
                _core.sync(dev)
                raise KeyboardInterrupt(str(e)) from None
            self.vol.write_back(x)
            _core.sync(dev)

    def get_res_norm(self):
        dev = self.proj.device()
        h = self._handle(dev)
        with torch.cuda.device(dev):
            self._bind_masks(h, dev)
            b = self.sino.on_device(dev)
            x = self.vol.on_device(dev)
            r = ctypes.c_double()
            _lib.call("astra_amd_algo_residual_norm", h, _core.ptr(b), _core.ptr(x), ctypes.byref(r),
                      _core.stream_ptr(dev))
        return float(r.value)

    def release(self):
        for dev, h in self._handles.items():
            try:
                with torch.cuda.device(dev):
                    _lib.call("astra_amd_algo_destroy", h)
            except Exception:
                pass
        self._handles.clear()


class SIRT_CUDA(_Iterative):
    algo_type = _lib.ALGO_SIRT


class CGLS_CUDA(_Iterative):
    algo_type = _lib.ALGO_CGLS


class SIRT3D_CUDA(_Iterative):
    algo_type = _lib.ALGO_SIRT
    is3d = True


class CGLS3D_CUDA(_Iterative):
    algo_type = _lib.ALGO_CGLS
    is3d = True


class EM_CUDA(_Iterative):
    """CCudaEMAlgorithm (src/CudaEMAlgorithm.cpp:63-253): multiplicative
    updates; no mask options."""
    algo_type = _lib.ALGO_EM

    def __init__(self, cfg):
        super().__init__(cfg)
        if self.masks:
            raise AstraError("EM_CUDA: mask options are not supported")


class SART_CUDA(_Iterative):
    """CCudaSartAlgorithm (src/CudaSartAlgorithm.cpp:63-316): one angle per
    iteration.  ProjectionOrder 'sequential', 'custom' (ProjectionOrderList)
    or 'random' (default): a permutation of the angles, drawn here from a
    fixed-seed generator (the reference shuffles with the C library's
    rand()), so 'random' runs are reproducible but not the reference's
    sequence."""
    algo_type = _lib.ALGO_SART

    def __init__(self, cfg):
        super().__init__(cfg)
        if "SinogramMask" in self.masks:
            raise AstraError("SART_CUDA: Sinogram mask option is not supported.")
        self.options["Relaxation"] = float(_opt(cfg, "Relaxation", 1.0))
        na = self.proj.proj_shape[0]
        order = str(_opt(cfg, "ProjectionOrder", "random")).lower()
        if order == "sequential":
            self.order = list(range(na))
        elif order == "random":
            self.order = [int(i) for i in np.random.default_rng(0).permutation(na)]
        elif order == "custom":
            lst = _opt(cfg, "ProjectionOrderList")
            if lst is None:
                raise AstraError("SART_CUDA: ProjectionOrder 'custom' needs ProjectionOrderList")
            self.order = [int(i) for i in np.asarray(lst).ravel()]
        else:
            raise AstraError("Unknown ProjectionOrder")

    def _handle(self, dev):
        new = dev.index not in self._handles
        h = super()._handle(dev)
        if new:
            arr = np.ascontiguousarray(self.order, dtype=np.int32)
            _lib.call("astra_amd_algo_set_order", h, arr.ctypes.data_as(ctypes.c_void_p), len(arr))
        return h


TYPES = {c.__name__: c for c in (FP_CUDA, BP_CUDA, FBP_CUDA, SIRT_CUDA, CGLS_CUDA, EM_CUDA, SART_CUDA,
                                  FP3D_CUDA, BP3D_CUDA, FDK_CUDA, SIRT3D_CUDA, CGLS3D_CUDA)}
NOT_BUILT = ("FP", "BP", "FBP", "SIRT", "SART", "ART", "CGLS",
             "DART_MASK_CUDA", "DART_SMOOTHING_CUDA", "DART_MASK_CUDA3D",
             "DART_SMOOTHING_CUDA3D", "ROI_SELECT_CUDA", "DATA_OPERATION_CUDA")


class _Plugin:
    """An algorithm object of a registered Python plugin
    (python/astra/src/PythonPluginAlgorithm.cpp:98-125)."""

    def __init__(self, cls, cfg):
        self.instance = cls()
        self.instance.astra_init(cfg)

    def run(self, iterations=1):
        self.instance.run(int(iterations))

    def get_res_norm(self):
        raise AstraError("this algorithm has no residual norm")

    def description(self):
        return "plugin " + type(self.instance).__name__

    def release(self):
        pass


def create(config):
    t = config.get("type")
    if t in TYPES:
        return _reg.add(TYPES[t](config))
    from . import plugin
    cls = plugin.lookup(t)
    if cls is not None:
        return _reg.add(_Plugin(cls, config))
    if t in NOT_BUILT:
        raise AstraError("algorithm type '%s' is not part of this build" % t)
    raise AstraError("unknown algorithm type '%s'" % t)


def run(i, iterations=1):
    _reg.get(i).run(iterations)


def get_res_norm(i):
    return _reg.get(i).get_res_norm()


def delete(ids):
    _reg.delete(ids)


def clear():
    _reg.clear()


def info():
    return _reg.info()


def get_plugin_object(i):
    """The Python object of a plugin algorithm (python/astra/algorithm.py:69-77)."""
    a = _reg.get(i)
    if not isinstance(a, _Plugin):
        raise AstraError("Not a plugin algorithm")
    return a.instance
