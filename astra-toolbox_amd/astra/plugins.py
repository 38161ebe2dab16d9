"""Example plugins (python/astra/plugins/sirt.py, python/astra/plugins/cgls.py):
SIRT and CGLS written in numpy on ``OpTomo`` (the GPU projector as a linear
operator).  Register them with ``astra.plugin.register(astra.plugins.SIRTPlugin)``;
their algorithm types are then "SIRT-PLUGIN" and "CGLS-PLUGIN", with a
ProjectorId, ProjectionDataId and ReconstructionDataId in the configuration
(2D or 3D data objects held in host memory).
"""
import numpy as np

from . import data2d, data3d, plugin
from ._registry import data2d as _d2
from .optomo import OpTomo

__all__ = ["SIRTPlugin", "CGLSPlugin"]


def _data_module(i):
    return data2d if i in _d2.objects else data3d


class SIRTPlugin(plugin.base):
    """SIRT.

    Options:

    'Relaxation': relaxation factor (optional)
    'MinConstraint': constrain values to at least this (optional)
    'MaxConstraint': constrain values to at most this (optional)
    """

    astra_name = "SIRT-PLUGIN"

    def initialize(self, cfg, Relaxation=1, MinConstraint=None, MaxConstraint=None):
        self.W = OpTomo(cfg["ProjectorId"])
        self.vid, self.sid = cfg["ReconstructionDataId"], cfg["ProjectionDataId"]
        self.mod = _data_module(self.sid)
        self.lo, self.hi = MinConstraint, MaxConstraint
        # x += C W^T R (b - W x): R = 1 / row sums, C = relaxation / column
        # sums, each 0 where the sum is below 1e-6
        rows = self.W.FP(np.ones(self.W.vshape, np.float32))
        cols = self.W.BP(np.ones(self.W.sshape, np.float32))
        with np.errstate(divide="ignore"):
            self.R = np.where(rows > 1e-6, 1.0 / rows, 0.0).astype(np.float32)
            self.C = np.where(cols > 1e-6, float(Relaxation) / cols, 0.0).astype(np.float32)

    def run(self, its):
        v = self.mod.get_shared(self.vid)
        s = self.mod.get_shared(self.sid)
        for _ in range(its):
            r = (s - self.W.FP(v)) * self.R
            v += self.C * self.W.BP(r)
            if self.lo is not None or self.hi is not None:
                np.clip(v, self.lo, self.hi, out=v)


class CGLSPlugin(plugin.base):
    """CGLS (conjugate gradients on the normal equations W^T W x = W^T b),
    restarted at every run from the current reconstruction."""

    astra_name = "CGLS-PLUGIN"

    def initialize(self, cfg):
        self.W = OpTomo(cfg["ProjectorId"])
        self.vid, self.sid = cfg["ReconstructionDataId"], cfg["ProjectionDataId"]
        self.mod = _data_module(self.sid)

    def run(self, its):
        x = self.mod.get_shared(self.vid)
        b = self.mod.get_shared(self.sid)
        W = self.W
        r = b - W.FP(x)
        p = W.BP(r)
        gamma = float(np.vdot(p.astype(np.float64), p))
        for _ in range(its):
            if gamma == 0.0:
                break
            w = W.FP(p)
            alpha = gamma / float(np.vdot(w.astype(np.float64), w))
            x += np.float32(alpha) * p
            r -= np.float32(alpha) * w
            z = W.BP(r)
            g1 = float(np.vdot(z.astype(np.float64), z))
            p = z + np.float32(g1 / gamma) * p
            gamma = g1
