"""MATLAB-style calls (python/astra/matlab.py): each function takes the
operation name as its first argument, as the MATLAB interface does, and
forwards to the module of the same name — ``astra.m.data2d('create',
'-vol', vol_geom, 0)``, ``astra.m.algorithm('iterate', alg_id, 100)``
('iterate' is 'run').  Available after ``import astra`` as ``astra.m``.
"""
from . import algorithm as _algorithm
from . import astra as _astra
from . import data2d as _data2d
from . import data3d as _data3d
from . import matrix as _matrix
from . import projector as _projector
from . import projector3d as _projector3d
from .log import AstraError

__all__ = ["astra", "data2d", "data3d", "projector", "projector3d", "matrix", "algorithm"]


def _call(module, command, args):
    fn = getattr(module, command, None)
    if fn is None or command.startswith("_"):
        raise AstraError("unknown %s command '%s'" % (module.__name__.rsplit(".", 1)[-1], command))
    return fn(*args)


def astra(command, *args):
    """``astra.m.astra('use_cuda')`` and the other astra.astra calls."""
    return _call(_astra, command, args)


def data2d(command, *args):
    """``astra.m.data2d('create', type, geometry, data)`` etc."""
    return _call(_data2d, command, args)


def data3d(command, *args):
    """``astra.m.data3d('get', i)`` etc."""
    return _call(_data3d, command, args)


def projector(command, *args):
    """``astra.m.projector('volume_geometry', i)`` etc."""
    return _call(_projector, command, args)


def projector3d(command, *args):
    """``astra.m.projector3d('projection_geometry', i)`` etc."""
    return _call(_projector3d, command, args)


def matrix(command, *args):
    """``astra.m.matrix('delete', i)`` etc."""
    return _call(_matrix, command, args)


def algorithm(command, *args):
    """``astra.m.algorithm('run', i, 1000)``; 'iterate' is 'run'."""
    if command == "iterate":
        command = "run"
    return _call(_algorithm, command, args)
