"""ctypes binding of libastra_amd.so (the C ABI declared in include/astra_amd.h).

The library is built in-tree (``astra-toolbox_amd/lib/libastra_amd.so``).
There is no fallback: if the library is missing or fails to load, every
projector / algorithm call raises.
"""
import ctypes
import os
import threading

from .log import AstraError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get(
    "ASTRA_AMD_LIB", os.path.join(os.path.dirname(_HERE), "lib", "libastra_amd.so"))

OK = 0
ERR_ARG, ERR_HIP, ERR_UNSUPPORTED, ERR_ALLOC, ERR_INTERNAL, ERR_ABORTED = 1, 2, 3, 4, 5, 6
MODE_ADD, MODE_SET = 0, 1
GEOM_PAR, GEOM_FAN, GEOM_PAR3D, GEOM_CONE, GEOM_CYLCONE = 0, 1, 2, 3, 4
KER3D_DEFAULT, KER3D_SUM_SQUARE_WEIGHTS, KER3D_FDK_WEIGHTING, KER3D_2D_WEIGHTING, KER3D_MATCHED_BP = range(5)
ALGO_SIRT, ALGO_CGLS, ALGO_EM, ALGO_SART = 0, 1, 2, 3


class Params2D(ctypes.Structure):
    _fields_ = [("rays_per_det", ctypes.c_uint), ("rays_per_pixel_dim", ctypes.c_uint),
                ("output_scale", ctypes.c_float)]


class Params3D(ctypes.Structure):
    _fields_ = [("rays_per_det_dim", ctypes.c_uint), ("rays_per_voxel_dim", ctypes.c_uint),
                ("output_scale", ctypes.c_float), ("kernel", ctypes.c_int)]


_vp = ctypes.c_void_p
_u = ctypes.c_uint
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double
_sz = ctypes.c_size_t

# int (*)(void *user): non-zero stops an iterative run
ABORT_HOOK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)

# name -> argtypes (all functions return int status, except the two strings)
SIGNATURES = {
    "astra_amd_set_device": [_i],
    "astra_amd_stream_sync": [_vp],
    "astra_amd_geom2d_create": [_i, _u, _u, _u, _u, _vp, ctypes.POINTER(_vp)],
    "astra_amd_geom2d_destroy": [_vp],
    "astra_amd_geom3d_create": [_i, _u, _u, _u, _u, _u, _u, _vp, _vp, ctypes.POINTER(_vp)],
    "astra_amd_geom3d_destroy": [_vp],
    "astra_amd_fp2d": [_vp, _vp, _vp, ctypes.POINTER(Params2D), _i, _vp],
    "astra_amd_fp2d_noalloc": [_vp, _vp, _vp, ctypes.POINTER(Params2D), _i, _vp],
    "astra_amd_bp2d": [_vp, _vp, _vp, ctypes.POINTER(Params2D), _i, _vp],
    "astra_amd_bp2d_fbp_weighted": [_vp, _vp, _vp, ctypes.POINTER(Params2D), _i, _vp],
    "astra_amd_bp2d_matched": [_vp, _vp, _vp, ctypes.POINTER(Params2D), _i, _vp],
    "astra_amd_fp3d": [_vp, _vp, _vp, ctypes.POINTER(Params3D), _i, _vp],
    "astra_amd_bp3d": [_vp, _vp, _vp, ctypes.POINTER(Params3D), _i, _vp],
    "astra_amd_fp2d_pitched": [_vp, _vp, _sz, _vp, _sz, ctypes.POINTER(Params2D), _i, _vp],
    "astra_amd_bp2d_pitched": [_vp, _vp, _sz, _vp, _sz, ctypes.POINTER(Params2D), _i, _vp],
    "astra_amd_fp3d_pitched": [_vp, _vp, _sz, _vp, _sz, ctypes.POINTER(Params3D), _i, _vp],
    "astra_amd_bp3d_pitched": [_vp, _vp, _sz, _vp, _sz, ctypes.POINTER(Params3D), _i, _vp],
    "astra_amd_ramlak_spectrum": [_u, _f, _vp, _u],
    "astra_amd_filter_spectrum": [_i, _u, _f, _f, _vp, _u],
    "astra_amd_filter_create": [_u, _vp, _u, ctypes.POINTER(_vp)],
    "astra_amd_filter_create_custom": [_u, _u, _i, _vp, _u, ctypes.POINTER(_vp)],
    "astra_amd_filter_destroy": [_vp],
    "astra_amd_filter_rows": [_vp, _vp, _sz, _vp],
    "astra_amd_fdk_preweight": [_vp, _vp, _vp],
    "astra_amd_fdk": [_vp, _vp, _vp, _f, _i, _vp, _i, _vp],
    "astra_amd_fdk_filter": [_vp, _vp, _i, _vp, _vp],
    "astra_amd_fbp2d": [_vp, _vp, _vp, _f, _i, _vp, _i, _vp],
    "astra_amd_dot": [_vp, _vp, _sz, ctypes.POINTER(_d), _vp],
    "astra_amd_algo2d_create": [_i, _vp, ctypes.POINTER(Params2D), ctypes.POINTER(_vp)],
    "astra_amd_algo3d_create": [_i, _vp, ctypes.POINTER(Params3D), ctypes.POINTER(_vp)],
    "astra_amd_algo_set_option": [_vp, ctypes.c_char_p, _d],
    "astra_amd_algo_set_mask": [_vp, ctypes.c_char_p, _vp],
    "astra_amd_algo_set_order": [_vp, _vp, _u],
    "astra_amd_algo_run": [_vp, _vp, _vp, _i, _vp],
    "astra_amd_set_abort_hook": [ABORT_HOOK, _vp],
    "astra_amd_algo_residual_norm": [_vp, _vp, _vp, ctypes.POINTER(_d), _vp],
    "astra_amd_algo_destroy": [_vp],
    "astra_amd_batch2d_create": [_u, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)],
    "astra_amd_batch2d_offsets": [_vp, _vp, _vp],
    "astra_amd_batch2d_fp": [_vp, _vp, _vp, _vp],
    "astra_amd_batch2d_sirt": [_vp, _vp, _vp, _i, _f, _f, _f, _i, _vp],
    "astra_amd_batch2d_destroy": [_vp],
    "astra_amd_profile_enable": [_i],
    "astra_amd_profile_read": [ctypes.c_char_p, ctypes.POINTER(_d), ctypes.POINTER(_d),
                               ctypes.POINTER(ctypes.c_ulonglong)],
    "astra_amd_calibrate_gather": [_i, _u, ctypes.POINTER(_d), _vp],
    "astra_amd_scratch_usage": [ctypes.POINTER(_sz), ctypes.POINTER(_sz), _i],
}
STRINGS = ["astra_amd_last_error", "astra_amd_version"]

_lock = threading.Lock()
_lib = None


def load():
    """Load the native library (once); raise AstraError if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise AstraError(
                "libastra_amd.so is not built (%s); run __graft_entry__.build() or "
                "python astra-toolbox_amd/build_lib.py" % LIB_PATH)
        # torch ships its own libamdhip64.so.7; loading it first makes the
        # library bind to that same HIP runtime (same SONAME), so torch
        # streams and allocations are valid handles inside the library.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        for name in STRINGS:
            fn = getattr(lib, name)
            fn.argtypes = []
            fn.restype = ctypes.c_char_p
        _lib = lib
        return lib


class Aborted(AstraError):
    """An iterative run stopped by the abort hook (status ERR_ABORTED)."""


def call(name, *args):
    """Call an entry point and turn a non-zero status into AstraError."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != OK:
        msg = lib.astra_amd_last_error().decode(errors="replace")
        cls = Aborted if st == ERR_ABORTED else AstraError
        raise cls("%s failed (status %d): %s" % (name, st, msg))
    return st


def version():
    return load().astra_amd_version().decode()


def profile_enable(on=True):
    """Start (clearing previous records) or stop per-kernel event timing."""
    call("astra_amd_profile_enable", 1 if on else 0)


def profile_read(kernel):
    """(total_ms, work_units, launches) recorded for a kernel name."""
    ms, units, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_ulonglong()
    call("astra_amd_profile_read", kernel.encode(), ctypes.byref(ms), ctypes.byref(units),
         ctypes.byref(n))
    return ms.value, units.value, n.value


def calibrate_gather(mode, iterations=4096):
    """Lane-loads per second of the gather calibration kernel on the
    current device (mode 0: 16-byte loads with the 3D taps' lane pattern,
    1: 16-byte coalesced, 2: 8-byte 2D taps at 4-byte alignment, 3: 16-byte
    loads of two 2D taps at 4-byte alignment, 4: the same at 8-byte
    alignment)."""
    out = ctypes.c_double()
    call("astra_amd_calibrate_gather", int(mode), int(iterations), ctypes.byref(out), None)
    return out.value


def scratch_usage(reset_peak=False):
    """(bytes in use, high-water mark) of the library's stream-ordered device
    scratch pool on the current device; reset_peak resets the mark after
    reading."""
    cur, peak = ctypes.c_size_t(), ctypes.c_size_t()
    call("astra_amd_scratch_usage", ctypes.byref(cur), ctypes.byref(peak), 1 if reset_peak else 0)
    return cur.value, peak.value
