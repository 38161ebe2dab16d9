"""The projector as a scipy LinearOperator (OpTomo, python/astra/optomo.py).

``W = astra.OpTomo(projector_id)`` behaves as the system matrix: ``W * x``
and ``W.matvec`` forward project a volume (flattened or in its array shape)
to a flattened sinogram, ``W.T * y`` and ``W.rmatvec`` back project, so
scipy's iterative solvers (lsqr, cg on the normal equations, ...) run on
the GPU projector.  ``FP`` / ``BP`` take an ``out`` array of the data
shape, and ``reconstruct`` runs one of the GPU algorithms on a sinogram.

Every product is one GPU FP or BP of the projector (the direct DLPack path
of astra.experimental: host arrays are staged through the projector's GPU).
"""
import numpy as np
import scipy.sparse.linalg

from . import algorithm, creators, data2d, data3d, experimental, geometry as G
from ._registry import projectors2d as _p2, projectors3d as _p3
from .log import AstraError

__all__ = ["OpTomo"]


def _as_data(a, shape):
    """float32, C-contiguous, in `shape` (a flattened vector is reshaped)."""
    a = np.asarray(a)
    if a.ndim == 1:
        a = a.reshape(shape)
    if a.shape != tuple(shape):
        raise AstraError("array of shape %s where %s was expected" % (a.shape, tuple(shape)))
    return np.ascontiguousarray(a, dtype=np.float32)


class OpTomo(scipy.sparse.linalg.LinearOperator):
    """System matrix of a 2D ('cuda') or 3D ('cuda3d') projector: shape
    (sinogram size, volume size), float32."""

    def __init__(self, proj_id):
        if proj_id in _p2.objects:
            self._p = _p2.get(proj_id)
            self.data_mod = data2d
        elif proj_id in _p3.objects:
            self._p = _p3.get(proj_id)
            self.data_mod = data3d
        else:
            raise AstraError("Projector not found: %r" % (proj_id,))
        self.proj_id = proj_id
        self.vg = G.config_form(self._p.vol_geom)
        self.pg = self._p.proj_geom
        self.vshape = tuple(self._p.vol_shape)
        self.sshape = tuple(self._p.proj_shape)
        self.vsize = int(np.prod(self.vshape))
        self.ssize = int(np.prod(self.sshape))
        super().__init__(np.float32, (self.ssize, self.vsize))
        self.transposeOpTomo = OpTomoTranspose(self)

    # -- scipy LinearOperator
    def _matvec(self, v):
        return self.FP(v).ravel()

    def _rmatvec(self, s):
        return self.BP(s).ravel()

    def rmatvec(self, s):
        return self._rmatvec(np.asarray(s).ravel())

    def _transpose(self):
        return self.transposeOpTomo

    _adjoint = _transpose

    def __mul__(self, v):
        if isinstance(v, np.ndarray) and v.shape == self.vshape:
            return self._matvec(v)
        return scipy.sparse.linalg.LinearOperator.__mul__(self, v)

    # -- projections
    def FP(self, v, out=None):
        """Forward projection of `v` (volume shape or flattened) into `out`
        (sinogram shape, float32, contiguous; allocated when None)."""
        v = _as_data(v, self.vshape)
        if out is None:
            out = np.zeros(self.sshape, dtype=np.float32)
        if self._p.is3d:
            experimental.direct_FP3D(self.proj_id, v, out)
        else:
            experimental.direct_FP2D(self.proj_id, v, out)
        return out

    def BP(self, s, out=None):
        """Back projection of `s` (sinogram shape or flattened) into `out`
        (volume shape, float32, contiguous; allocated when None)."""
        s = _as_data(s, self.sshape)
        if out is None:
            out = np.zeros(self.vshape, dtype=np.float32)
        if self._p.is3d:
            experimental.direct_BP3D(self.proj_id, out, s)
        else:
            experimental.direct_BP2D(self.proj_id, out, s)
        return out

    def reconstruct(self, method, s, iterations=1, extraOptions=None):
        """Run algorithm `method` (e.g. 'SIRT_CUDA', 'FBP_CUDA', 'CGLS3D_CUDA')
        for `iterations` on sinogram `s` from a zero volume; returns the
        volume.  `extraOptions` become the algorithm's options."""
        s = _as_data(s, self.sshape)
        v = np.zeros(self.vshape, dtype=np.float32)
        sid = self.data_mod.link("-sino", self.pg, s)
        vid = self.data_mod.link("-vol", self.vg, v)
        try:
            cfg = creators.astra_dict(method)
            cfg.update(ProjectionDataId=sid, ReconstructionDataId=vid, ProjectorId=self.proj_id)
            cfg["option"] = dict(extraOptions or {})
            aid = algorithm.create(cfg)
            try:
                algorithm.run(aid, iterations)
            finally:
                algorithm.delete(aid)
        finally:
            self.data_mod.delete([vid, sid])
        return v


class OpTomoTranspose(scipy.sparse.linalg.LinearOperator):
    """``OpTomo.T``: back projection as the operator, forward projection as
    its transpose."""

    def __init__(self, parent):
        self.parent = parent
        super().__init__(np.float32, (parent.shape[1], parent.shape[0]))

    def _matvec(self, s):
        return self.parent._rmatvec(s)

    def _rmatvec(self, v):
        return self.parent._matvec(v)

    def rmatvec(self, v):
        return self._rmatvec(np.asarray(v).ravel())

    def _transpose(self):
        return self.parent

    _adjoint = _transpose

    def __mul__(self, s):
        if isinstance(s, np.ndarray) and s.shape == self.parent.sshape:
            return self._matvec(s)
        return scipy.sparse.linalg.LinearOperator.__mul__(self, s)
