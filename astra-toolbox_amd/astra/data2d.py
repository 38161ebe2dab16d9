"""2D data objects (python/astra/data2d.py:28-160)."""
import numpy as np

from . import _core
from . import geometry as G
from ._registry import data2d as _reg
from .log import AstraError

_KINDS = ("-vol", "-sino")


def _shape(datatype, geometry):
    if datatype == "-vol":
        if G.is_3d_vol(geometry):
            raise AstraError("2D data object needs a 2D volume geometry")
        return G.vol_shape(geometry)
    if datatype == "-sino":
        if G.is_3d_proj(geometry):
            raise AstraError("2D data object needs a 2D projection geometry")
        return G.proj_shape(geometry)
    raise AstraError("unknown 2D data type '%s'" % datatype)


def create(datatype, geometry, data=None):
    """Create a 2D object; `data` is a scalar fill value or an array."""
    shape = _shape(datatype, geometry)
    return _reg.add(_core.new_data(datatype, geometry, shape, data))


def link(datatype, geometry, data):
    """Create a 2D object sharing the memory of a DLPack array (no copy)."""
    shape = _shape(datatype, geometry)
    return _reg.add(_core.link_data(datatype, geometry, shape, data))


def store(i, data):
    _reg.get(i).store(data)


def get_geometry(i):
    """The object's geometry in the reference's configuration form (a
    volume's window under 'options')."""
    return G.config_form(_reg.get(i).geom)


def change_geometry(i, geom):
    d = _reg.get(i)
    if _shape(d.kind, geom) != d.shape:
        raise AstraError("new geometry has a different data shape")
    d.geom = geom


def get(i):
    return _reg.get(i).get()


def get_shared(i):
    return _reg.get(i).numpy_view()


def get_single(i):
    return get(i).astype(np.float32)


def delete(ids):
    _reg.delete(ids)


def clear():
    _reg.clear()


def info():
    return _reg.info()


def shepp_logan(geometry, modified=True, returnData=True):
    """A volume data object holding the Shepp-Logan phantom on the
    geometry's grid (data2d.shepp_logan, python/astra/data2d.py); modified=False
    gives the original densities.  Returns the id, or (id, data)."""
    from .phantom import shepp_logan_2d
    i = create("-vol", geometry)
    _reg.get(i).store(shepp_logan_2d(*G.vol_shape(geometry), modified=bool(modified)))
    return (i, get(i)) if returnData else i


def dimensions(i):
    return _reg.get(i).shape
