"""3D projector objects (python/astra/projector3d.py:28-128); type 'cuda3d'
(CCudaProjector3D, src/CudaProjector3D.cpp:83-124)."""
from . import _core
from . import geometry as G
from ._registry import projectors3d as _reg
from .log import AstraError


def create(config):
    t = config.get("type")
    if t != "cuda3d":
        raise AstraError("unknown 3D projector type '%s'" % t)
    opts = G.options_of(config)
    p = _core.Projector("cuda3d", config["ProjectionGeometry"], config["VolumeGeometry"], opts)
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
    return True


def weights_single_ray(i, projection_index, detector_index):
    raise AstraError("explicit projection weights are not available for GPU projectors")


def weights_projection(i, projection_index):
    raise AstraError("explicit projection weights are not available for GPU projectors")


def splat(i, row, col):
    raise AstraError("splat is not available for GPU projectors")


def direct_FP(projector_id, volume, *, out):
    from .experimental import direct_FP3D
    direct_FP3D(projector_id, volume, out)
    return out


def direct_BP(projector_id, projection, *, out):
    from .experimental import direct_BP3D
    direct_BP3D(projector_id, out, projection)
    return out
