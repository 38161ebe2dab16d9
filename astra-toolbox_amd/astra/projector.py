"""2D projector objects (python/astra/projector.py:28-138).

Only the GPU projector type 'cuda' (CCudaProjector2D) is built; the CPU
projector types of the reference ('line', 'linear', 'strip', ...) are
outside this build's hot path and raise.
"""
from . import _core
from . import geometry as G
from ._registry import projectors2d as _reg
from .log import AstraError

CPU_TYPES = ("line", "linear", "strip", "line_fanflat", "strip_fanflat", "distance_driven",
             "sparse_matrix", "blob")


def create(config):
    t = config.get("type")
    if t in CPU_TYPES:
        raise AstraError("CPU projector type '%s' is not part of this build; use 'cuda'" % t)
    if t != "cuda":
        raise AstraError("unknown 2D projector type '%s'" % t)
    opts = G.options_of(config)
    p = _core.Projector("cuda", config["ProjectionGeometry"], config["VolumeGeometry"], opts)
    return _reg.add(p)


def get(i):
    return _reg.get(i)


def delete(ids):
    _reg.delete(ids)


def clear():
    _reg.clear()


def info():
    return _reg.info()


def projection_geometry(i):
    return _reg.get(i).proj_geom


def volume_geometry(i):
    """The projector's volume geometry in configuration form (window under
    'options'), as the reference returns it."""
    return G.config_form(_reg.get(i).vol_geom)


def is_cuda(i):
    return _reg.get(i).type == "cuda"


def weights_single_ray(i, projection_index, detector_index):
    raise AstraError("explicit projection weights are not available for GPU projectors")


def weights_projection(i, projection_index):
    raise AstraError("explicit projection weights are not available for GPU projectors")


def splat(i, row, col):
    raise AstraError("splat is not available for GPU projectors")


def matrix(i):
    raise AstraError("sparse matrices are not available for GPU projectors")


def direct_FP(projector_id, volume, *, out):
    """Forward projection between pre-allocated DLPack tensors (CPU or GPU)."""
    from .experimental import direct_FP2D
    direct_FP2D(projector_id, volume, out)
    return out


def direct_BP(projector_id, projection, *, out):
    """Back projection between pre-allocated DLPack tensors (CPU or GPU)."""
    from .experimental import direct_BP2D
    direct_BP2D(projector_id, out, projection)
    return out
