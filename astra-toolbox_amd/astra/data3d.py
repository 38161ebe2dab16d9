"""3D data objects (python/astra/data3d.py:30-175)."""
import numpy as np

from . import _core
from . import geometry as G
from ._registry import data3d as _reg
from .log import AstraError


def _shape(datatype, geometry):
    if datatype == "-vol":
        if not G.is_3d_vol(geometry):
            raise AstraError("3D data object needs a 3D volume geometry")
        return G.vol_shape(geometry)
    if datatype in ("-sino", "-proj3d"):
        if not G.is_3d_proj(geometry):
            raise AstraError("3D data object needs a 3D projection geometry")
        return G.proj_shape(geometry)
    raise AstraError("unknown 3D data type '%s'" % datatype)


def create(datatype, geometry, data=None):
    shape = _shape(datatype, geometry)
    return _reg.add(_core.new_data("-vol" if datatype == "-vol" else "-sino", geometry, shape, data))


def link(datatype, geometry, data):
    shape = _shape(datatype, geometry)
    return _reg.add(_core.link_data("-vol" if datatype == "-vol" else "-sino", geometry, shape, data))


def get(i):
    return _reg.get(i).get()


def get_shared(i):
    return _reg.get(i).numpy_view()


def get_single(i):
    return get(i).astype(np.float32)


def store(i, data):
    _reg.get(i).store(data)


def get_geometry(i):
    """The object's geometry in the reference's configuration form (a
    volume's window under 'options')."""
    return G.config_form(_reg.get(i).geom)


def change_geometry(i, geometry):
    d = _reg.get(i)
    if _shape(d.kind if d.kind == "-vol" else "-sino", geometry) != d.shape:
        raise AstraError("new geometry has a different data shape")
    d.geom = geometry


def dimensions(i):
    return _reg.get(i).shape


def delete(ids):
    _reg.delete(ids)


def clear():
    _reg.clear()


def info():
    return _reg.info()


def shepp_logan(geometry, modified=True, returnData=True):
    """A volume data object holding the Shepp-Logan phantom on the
    geometry's grid (data3d.shepp_logan, python/astra/data3d.py); modified=False
    gives the original densities.  Returns the id, or (id, data)."""
    from .phantom import shepp_logan_3d
    i = create("-vol", geometry)
    _reg.get(i).store(shepp_logan_3d(*G.vol_shape(geometry), modified=bool(modified)))
    return (i, get(i)) if returnData else i
