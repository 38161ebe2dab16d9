"""ctypes access to the CPU checkers built by oracle/Makefile.

Test infrastructure only (see ``oracle/__init__.py``):
* ``port()``      -> the C restatement (oracle/cpu_linear.c)
* ``reference()`` -> the reference's own CPU library (oracle/_ref), or None
  when it was not built (it can only be built where the reference sources
  are present).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PORT_LIB = os.path.join(HERE, "_build", "libcpu_linear.so")
REF_LIB = os.path.join(HERE, "_ref", "libastra_ref_cpu.so")

_fp = ctypes.POINTER(ctypes.c_float)


class LinearCPU:
    """parallel-beam 'linear' projector FP / BP / SIRT through one library."""

    def __init__(self, path, prefix):
        self.path = path
        self.lib = ctypes.CDLL(path)
        i, f = ctypes.c_int, ctypes.c_float
        self._fp = getattr(self.lib, prefix + "_linear_fp")
        self._bp = getattr(self.lib, prefix + "_linear_bp")
        self._sirt = getattr(self.lib, prefix + "_linear_sirt")
        self._fp.argtypes = [i, i, i, i, f, _fp, _fp, _fp]
        self._bp.argtypes = [i, i, i, i, f, _fp, _fp, _fp]
        if prefix == "cpu":
            self._sirt.argtypes = [i, i, i, i, f, _fp, _fp, _fp, i, f]
        else:
            self._sirt.argtypes = [i, i, i, i, f, _fp, _fp, _fp, i]
        self.prefix = prefix
        for fn in (self._fp, self._bp, self._sirt):
            fn.restype = ctypes.c_int

    @staticmethod
    def _p(a):
        return a.ctypes.data_as(_fp)

    def fp(self, vol, angles, n_det, det_width=1.0):
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        ang = np.ascontiguousarray(angles, dtype=np.float32)
        out = np.zeros((len(ang), n_det), dtype=np.float32)
        r = self._fp(vol.shape[0], vol.shape[1], len(ang), n_det, det_width, self._p(ang),
                     self._p(vol), self._p(out))
        assert r == 0
        return out

    def bp(self, sino, angles, shape, det_width=1.0):
        sino = np.ascontiguousarray(sino, dtype=np.float32)
        ang = np.ascontiguousarray(angles, dtype=np.float32)
        out = np.zeros(shape, dtype=np.float32)
        r = self._bp(shape[0], shape[1], len(ang), sino.shape[1], det_width, self._p(ang),
                     self._p(sino), self._p(out))
        assert r == 0
        return out

    def sirt(self, sino, angles, shape, iters, det_width=1.0, x0=None):
        sino = np.ascontiguousarray(sino, dtype=np.float32)
        ang = np.ascontiguousarray(angles, dtype=np.float32)
        out = np.zeros(shape, dtype=np.float32) if x0 is None else np.array(x0, dtype=np.float32)
        args = [shape[0], shape[1], len(ang), sino.shape[1], det_width, self._p(ang),
                self._p(sino), self._p(out), iters]
        if self.prefix == "cpu":
            args.append(1.0)
        assert self._sirt(*args) == 0
        return out


def port():
    return LinearCPU(PORT_LIB, "cpu")


def reference():
    if not os.path.exists(REF_LIB):
        return None
    return LinearCPU(REF_LIB, "ref")


def reference_shepp_logan(shape, modified=True):
    """The reference's generateSheppLogan / generateSheppLogan3D
    (src/SheppLogan.cpp:187-246) through oracle/_ref, or None when it was
    not built.  shape: (rows, cols) or (slices, rows, cols)."""
    if not os.path.exists(REF_LIB):
        return None
    lib = ctypes.CDLL(REF_LIB)
    out = np.zeros(shape, dtype=np.float32)
    if len(shape) == 2:
        r = lib.ref_shepp_logan_2d(int(shape[0]), int(shape[1]), int(bool(modified)), out.ctypes.data_as(_fp))
    else:
        r = lib.ref_shepp_logan_3d(int(shape[0]), int(shape[1]), int(shape[2]), int(bool(modified)),
                                   out.ctypes.data_as(_fp))
    assert r == 0
    return out


def reference_strip():
    """(fp, bp) of the reference's parallel 'strip' projector
    (include/astra/ParallelBeamStripKernelProjector2D.inl) through
    oracle/_ref, or None when it was not built."""
    if not os.path.exists(REF_LIB):
        return None
    lib = ctypes.CDLL(REF_LIB)
    i, f = ctypes.c_int, ctypes.c_float
    for fn in (lib.ref_strip_fp, lib.ref_strip_bp):
        fn.argtypes = [i, i, i, i, f, _fp, _fp, _fp]
        fn.restype = ctypes.c_int

    def fp(vol, angles, n_det, det_width=1.0):
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        ang = np.ascontiguousarray(angles, dtype=np.float32)
        out = np.zeros((len(ang), n_det), dtype=np.float32)
        assert lib.ref_strip_fp(vol.shape[0], vol.shape[1], len(ang), n_det, det_width,
                                ang.ctypes.data_as(_fp), vol.ctypes.data_as(_fp), out.ctypes.data_as(_fp)) == 0
        return out

    def bp(sino, angles, shape, det_width=1.0):
        sino = np.ascontiguousarray(sino, dtype=np.float32)
        ang = np.ascontiguousarray(angles, dtype=np.float32)
        out = np.zeros(shape, dtype=np.float32)
        assert lib.ref_strip_bp(shape[0], shape[1], len(ang), sino.shape[1], det_width,
                                ang.ctypes.data_as(_fp), sino.ctypes.data_as(_fp), out.ctypes.data_as(_fp)) == 0
        return out
    return fp, bp
