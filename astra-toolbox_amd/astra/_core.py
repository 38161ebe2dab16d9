"""Projector objects, data objects and the calls into the native library.

Roles in the reference: CCudaProjector2D / CCudaProjector3D
(src/CudaProjector2D.cpp, src/CudaProjector3D.cpp), CFloat32Data2D/3D with
host or GPU storage (include/astra/Data.h:40-120), and the single-GPU,
unchunked job execution of CCompositeGeometryManager
(src/CompositeGeometryManager.cpp:1115-1280).  Device memory and streams come
from torch; all arithmetic runs in libastra_amd.
"""
import ctypes
import threading

import numpy as np

from . import _lib
from . import geometry as G
from .log import AstraError

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the stack
    torch = None

_state = threading.local()
_default_gpu = [0]


def set_gpu_index(idx, memory=0):
    """astra.set_gpu_index (python/astra/astra_c.pyx:84-118).  A list picks its
    first entry: one device per process here.  `memory` (bytes, 0 = the
    device's free memory) bounds the GPU memory of a 3D FP / BP / FDK job on
    host data; larger jobs are split (astra.composite)."""
    if memory != 0 and memory < 1024 * 1024:
        raise AstraError("Setting GPU memory lower than 1MB is not supported")
    if isinstance(idx, (list, tuple)):
        if not idx:
            raise AstraError("set_gpu_index: empty GPU list")
        idx = idx[0]
    _default_gpu[0] = int(idx)
    _memory_limit[0] = int(memory)


_memory_limit = [0]


def gpu_memory_limit():
    """The job memory limit set by set_gpu_index (0 = not set)."""
    return _memory_limit[0]


def device(index=None):
    if torch is None or not torch.cuda.is_available():
        raise AstraError("no GPU available: the ASTRA GPU path needs a ROCm device")
    i = _default_gpu[0] if index is None else int(index)
    if i >= torch.cuda.device_count():
        raise AstraError("GPU index %d out of range" % i)
    return torch.device("cuda", i)


def stream_ptr(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def sync(dev):
    torch.cuda.current_stream(dev).synchronize()


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


# ------------------------------------------------------------- projectors -
class Projector:
    """A GPU projector: geometry dictionaries plus the options of
    CCudaProjector2D/3D (ProjectionKernel, VoxelSuperSampling,
    DetectorSuperSampling, GPUindex)."""

    def description(self):
        """CCudaProjector2D / 3D::description (src/CudaProjector2D.cpp:109-112,
        src/CudaProjector3D.cpp:148-151)."""
        return "" if self.type == "cuda3d" else "cuda"

    def __init__(self, ptype, proj_geom, vol_geom, options=None):
        options = dict(options or {})
        self.type = ptype
        self.proj_geom = proj_geom
        self.vol_geom = vol_geom
        self.options = options
        self.is3d = ptype == "cuda3d"
        if self.is3d != G.is_3d_proj(proj_geom) or self.is3d != G.is_3d_vol(vol_geom):
            raise AstraError("projector type '%s' does not match the geometry dimensions" % ptype)
        gi = options.get("GPUindex", options.get("GPUIndex", None))
        if isinstance(gi, (list, tuple)):
            gi = gi[0] if gi else None
        self.gpu_index = None if gi is None else int(gi)
        self.detector_ss = int(options.get("DetectorSuperSampling", 1))
        self.voxel_ss = int(options.get("VoxelSuperSampling", 1))
        kern = options.get("ProjectionKernel", "default")
        kmap = {"default": _lib.KER3D_DEFAULT, "sum_square_weights": _lib.KER3D_SUM_SQUARE_WEIGHTS,
                "2d_weighting": _lib.KER3D_2D_WEIGHTING, "matched_bp": _lib.KER3D_MATCHED_BP}
        if kern not in kmap:
            raise AstraError("Unknown ProjectionKernel '%s'" % kern)
        self.kernel = kmap[kern]
        self.vol_shape = G.vol_shape(vol_geom)
        self.proj_shape = G.proj_shape(proj_geom)
        if self.is3d:
            self.kind, vecs = G.vectors_3d(proj_geom)
            self.vecs, self.vol_scale = G.normalise_3d(self.kind, vecs, vol_geom)
            self.out_scale = 1.0
        else:
            self.kind, vecs = G.vectors_2d(proj_geom)
            self.vecs, self.out_scale = G.normalise_2d(self.kind, vecs, vol_geom)
        self._handles = {}
        self._lock = threading.Lock()

    def device(self):
        return device(self.gpu_index)

    def handle(self, dev):
        """Native geometry handle on `dev` (created once per device)."""
        key = dev.index
        with self._lock:
            h = self._handles.get(key)
            if h is not None:
                return h
            with torch.cuda.device(dev):
                _lib.call("astra_amd_set_device", key)
                out = ctypes.c_void_p()
                if self.is3d:
                    nz, ny, nx = self.vol_shape
                    nv, na, nu = self.proj_shape
                    v = np.ascontiguousarray(self.vecs, dtype=np.float64)
                    vs = (ctypes.c_float * 3)(*self.vol_scale)
                    kind = {"cone": _lib.GEOM_CONE, "par3d": _lib.GEOM_PAR3D,
                            "cylcone": _lib.GEOM_CYLCONE}[self.kind]
                    _lib.call("astra_amd_geom3d_create", kind, nx, ny, nz, na, nu, nv,
                              v.ctypes.data_as(ctypes.c_void_p), ctypes.cast(vs, ctypes.c_void_p),
                              ctypes.byref(out))
                else:
                    h_, w_ = self.vol_shape
                    na, nd = self.proj_shape
                    v = np.ascontiguousarray(self.vecs, dtype=np.float32)
                    kind = _lib.GEOM_FAN if self.kind == "fan" else _lib.GEOM_PAR
                    _lib.call("astra_amd_geom2d_create", kind, w_, h_, na, nd,
                              v.ctypes.data_as(ctypes.c_void_p), ctypes.byref(out))
            h = out.value
            self._handles[key] = h
            return h

    def params2d(self, scale=1.0):
        return _lib.Params2D(self.detector_ss, self.voxel_ss, float(self.out_scale * scale))

    def params3d(self, scale=1.0, kernel=None):
        k = self.kernel if kernel is None else kernel
        return _lib.Params3D(self.detector_ss, self.voxel_ss, float(self.out_scale * scale), k)

    def release(self):
        with self._lock:
            for dev, h in self._handles.items():
                try:
                    with torch.cuda.device(dev):
                        _lib.call("astra_amd_set_device", dev)
                        _lib.call("astra_amd_geom3d_destroy" if self.is3d else
                                  "astra_amd_geom2d_destroy", ctypes.c_void_p(h))
                except Exception:
                    pass
            self._handles.clear()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def projector_for(proj_geom, vol_geom, options=None):
    """Anonymous GPU projector for a geometry pair (the reference's algorithms
    build one when no ProjectorId is given)."""
    ptype = "cuda3d" if G.is_3d_proj(proj_geom) else "cuda"
    return Projector(ptype, proj_geom, vol_geom, options)


# ---------------------------------------------------------- native calls --
def _check_tensor(t, shape, what):
    if t.dtype != torch.float32:
        raise AstraError("%s must be float32" % what)
    if tuple(t.shape) != tuple(shape):
        raise AstraError("%s has shape %s, expected %s" % (what, tuple(t.shape), tuple(shape)))
    if not t.is_contiguous():
        raise AstraError("%s must be contiguous" % what)
    if t.device.type != "cuda":
        raise AstraError("%s must be a GPU tensor" % what)


def fp(proj, vol_t, proj_t, mode=_lib.MODE_SET, scale=1.0):
    """proj_t (=|+=) scale * A vol_t on the tensors' device (astraCUDA[3d]::FP)."""
    dev = vol_t.device
    _check_tensor(vol_t, proj.vol_shape, "volume")
    _check_tensor(proj_t, proj.proj_shape, "projection data")
    h = ctypes.c_void_p(proj.handle(dev))
    with torch.cuda.device(dev):
        s = stream_ptr(dev)
        if proj.is3d:
            p = proj.params3d(scale)
            _lib.call("astra_amd_fp3d", h, ptr(vol_t), ptr(proj_t), ctypes.byref(p), mode, s)
        else:
            p = proj.params2d(scale)
            _lib.call("astra_amd_fp2d", h, ptr(vol_t), ptr(proj_t), ctypes.byref(p), mode, s)


def bp(proj, proj_t, vol_t, mode=_lib.MODE_SET, scale=1.0, kernel=None, fbp_weighted=False):
    """vol_t (=|+=) scale * A^T proj_t (astraCUDA[3d]::BP)."""
    dev = vol_t.device
    _check_tensor(vol_t, proj.vol_shape, "volume")
    _check_tensor(proj_t, proj.proj_shape, "projection data")
    h = ctypes.c_void_p(proj.handle(dev))
    with torch.cuda.device(dev):
        s = stream_ptr(dev)
        if proj.is3d:
            p = proj.params3d(scale, kernel)
            _lib.call("astra_amd_bp3d", h, ptr(proj_t), ptr(vol_t), ctypes.byref(p), mode, s)
        else:
            p = proj.params2d(scale)
            if kernel == _lib.KER3D_MATCHED_BP:
                name = "astra_amd_bp2d_matched"
            elif fbp_weighted:
                name = "astra_amd_bp2d_fbp_weighted"
            else:
                name = "astra_amd_bp2d"
            _lib.call(name, h, ptr(proj_t), ptr(vol_t), ctypes.byref(p), mode, s)


# ---------------------------------------------------------------- data ----
class Data:
    """A 2D/3D float32 data object with host (numpy) or linked storage.

    Linked storage is any DLPack producer (numpy, torch CPU/GPU, ...): it is
    wrapped without a copy (python/astra/src/dlpack.cpp:41-304,
    cuda/dlpack_support.cu:38-100)."""

    def __init__(self, kind, geom, shape, host=None, linked=None):
        self.kind = kind
        self.geom = geom
        self.shape = tuple(shape)
        self.host = host
        self.linked = linked

    def description(self):
        """CData2D / CData3D::description (src/Data2D.cpp:34-41,
        src/Data3D.cpp:34-41): the dimensions fastest axis first."""
        what = "volume" if self.kind == "-vol" else "sinogram"
        return "%s %s data \t" % ("x".join(str(n) for n in reversed(self.shape)), what)

    # -- storage views
    def is_gpu(self):
        return self.linked is not None and self.linked.device.type == "cuda"

    def numpy_view(self):
        if self.linked is not None:
            if self.is_gpu():
                raise AstraError("data object is stored on the GPU")
            return self.linked.numpy()
        return self.host

    def get(self):
        if self.is_gpu():
            return self.linked.detach().cpu().numpy().copy()
        return np.array(self.numpy_view(), copy=True)

    def store(self, value):
        if np.isscalar(value):
            if self.linked is not None:
                self.linked.fill_(float(value))
            else:
                self.host.fill(value)
            return
        arr = np.asarray(value, dtype=np.float32)
        if arr.shape != self.shape:
            raise AstraError("data has shape %s, expected %s" % (arr.shape, self.shape))
        if self.linked is not None:
            self.linked.copy_(torch.from_numpy(np.ascontiguousarray(arr)))
        else:
            self.host[...] = arr

    # -- device access for one operation
    def on_device(self, dev, init=True):
        """Tensor on `dev` holding this object's data (the linked tensor
        itself when it already lives there and is contiguous; a pitched
        GPULink view goes through a contiguous copy and write_back)."""
        if self.is_gpu() and self.linked.device == dev:
            if self.linked.is_contiguous():
                return self.linked
            if init:
                return self.linked.contiguous()
            return torch.empty(self.shape, dtype=torch.float32, device=dev)
        src = self.linked if self.linked is not None else torch.from_numpy(self.host)
        if init:
            return src.to(dev, non_blocking=False).contiguous()
        return torch.empty(self.shape, dtype=torch.float32, device=dev)

    def write_back(self, t):
        """Copy an operation's result tensor back into this object."""
        if self.is_gpu() and t.data_ptr() == self.linked.data_ptr():
            return
        if self.linked is not None:
            self.linked.copy_(t)
        else:
            self.host[...] = t.cpu().numpy()


def new_data(kind, geom, shape, data=None):
    host = np.zeros(shape, dtype=np.float32)
    d = Data(kind, geom, shape, host=host)
    if data is not None:
        d.store(data)
    return d


# ---------------------------------------------------- raw device pointers -
class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", _DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", _DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    _fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p),
                ("deleter", ctypes.c_void_p)]


_KDL_ROCM = 10
_dl_keepalive = []  # the descriptors torch reads at free time (no deleter)


def _tensor_from_pointer(ptr, shape, row_pitch_elems, dev_index):
    """A torch view of device memory at `ptr` (float32, [z][y][x] with rows
    `row_pitch_elems` apart), through a DLPack capsule without a deleter:
    the caller owns the memory."""
    z, y, x = shape
    shp = (ctypes.c_int64 * 3)(z, y, x)
    strides = (ctypes.c_int64 * 3)(y * row_pitch_elems, row_pitch_elems, 1)
    mt = _DLManagedTensor()
    mt.dl_tensor.data = ctypes.c_void_p(int(ptr))
    mt.dl_tensor.device = _DLDevice(_KDL_ROCM, int(dev_index))
    mt.dl_tensor.ndim = 3
    mt.dl_tensor.dtype = _DLDataType(2, 32, 1)  # kDLFloat, 32 bits
    mt.dl_tensor.shape = shp
    mt.dl_tensor.strides = strides
    mt.dl_tensor.byte_offset = 0
    mt.manager_ctx = None
    mt.deleter = None
    _dl_keepalive.append((mt, shp, strides))
    new_capsule = ctypes.pythonapi.PyCapsule_New
    new_capsule.restype = ctypes.py_object
    new_capsule.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
    cap = new_capsule(ctypes.addressof(mt), b"dltensor", None)
    return torch.from_dlpack(cap)


def _link_gpulink(kind, geom, shape, link):
    """utils.pyx:305-317 / 335-347: shape check, then wrap the pointer."""
    data_shape = (int(link.z), int(link.y), int(link.x))
    if tuple(shape) != data_shape:
        raise ValueError("The dimensions of the data {} do not match those specified in the "
                         "geometry {}".format(data_shape, tuple(shape)))
    if int(link.pitch) % 4 or int(link.pitch) < 4 * int(link.x):
        raise ValueError("GPULink pitch must be a multiple of 4 bytes and at least 4 * x")
    if torch is None or not torch.cuda.is_available():
        raise AstraError("no GPU available: the ASTRA GPU path needs a ROCm device")
    dev = device()
    t = _tensor_from_pointer(link.ptr, data_shape, int(link.pitch) // 4, dev.index)
    return Data(kind, geom, shape, linked=t)


def link_data(kind, geom, shape, obj):
    """Wrap a DLPack producer without copying (python/astra/utils.pyx:237-349);
    3D data also takes a GPULink (a raw, possibly pitched, device pointer)."""
    from .pythonutils import GPULink
    if isinstance(obj, GPULink) and len(shape) == 3:
        return _link_gpulink(kind, geom, shape, obj)
    if not hasattr(obj, "__dlpack__"):
        raise TypeError("Object of type %s does not support the DLPack protocol" % type(obj).__name__)
    if isinstance(obj, np.ndarray) and not obj.flags.writeable:
        raise ValueError("Failed to link dlpack array: read-only arrays are not supported")
    try:
        t = torch.from_dlpack(obj)
    except Exception as e:  # noqa: BLE001 - map producer errors to ValueError
        raise ValueError("Failed to link dlpack array: %s" % e)
    if t.dtype != torch.float32:
        raise ValueError("Failed to link dlpack array: data type must be float32")
    squeezed = tuple(s for s in t.shape if s != 1)
    want = tuple(s for s in shape if s != 1)
    if squeezed != want:
        raise ValueError("Failed to link dlpack array: shape %s does not match geometry %s"
                         % (tuple(t.shape), tuple(shape)))
    if not t.is_contiguous():
        raise ValueError("Failed to link dlpack array: array must be C-contiguous")
    t = t.view(shape)
    return Data(kind, geom, shape, linked=t)
