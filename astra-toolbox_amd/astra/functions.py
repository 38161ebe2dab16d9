"""Helper functions (python/astra/functions.py:47-300)."""
import numpy as np

from . import _registry, geometry as G
from .log import AstraError


def clear():
    """Delete every object (functions.py:47-58)."""
    for r in (_registry.algorithms, _registry.data2d, _registry.data3d,
              _registry.projectors2d, _registry.projectors3d):
        r.clear()


def add_noise_to_sino(sinogram_in, I0, seed=None):
    """Poisson noise for I0 incident photons (functions.py:83-121):
    b = -max * log(Poisson(I0 exp(-p / max)) / I0), drawn from numpy's
    global legacy generator (seeded and restored when `seed` is given), one
    draw per element in C order, so a seed reproduces the reference's
    output.  A data id is updated in place."""
    if seed is not None:
        state = np.random.get_state()
        np.random.seed(seed)
    try:
        if isinstance(sinogram_in, np.ndarray):
            raw = sinogram_in
        else:
            from . import data2d
            raw = data2d.get(sinogram_in)
        mx = raw.max()
        ct = I0 * np.exp(-(raw / mx))
        counts = np.random.poisson(ct).astype(ct.dtype)
        with np.errstate(divide="ignore"):
            out = -mx * np.log(counts / I0)
        if not isinstance(sinogram_in, np.ndarray):
            data2d.store(sinogram_in, out)
        return out
    finally:
        if seed is not None:
            np.random.set_state(state)


def geom_2vec(proj_geom):
    return G.geom_2vec(proj_geom)


def geom_size(geom, dim=None):
    """Shape of the data that belongs to a geometry (functions.py:156-165)."""
    if "GridSliceCount" in geom or "GridRowCount" in geom:
        s = G.vol_shape(geom)
    else:
        s = G.proj_shape(geom)
    return s if dim is None else s[dim]


def move_vol_geom(geom, pos, is_relative=False):
    """Translate a volume geometry window (functions.py:123-154)."""
    import copy
    g = copy.deepcopy(geom)
    o = G.options_of(g)
    for k in G.OPTION_KEYS:
        g.pop(k, None)
    g["option"] = o
    axes = ["X", "Y", "Z"][: len(pos)]
    for ax, p in zip(axes, pos):
        w = o["WindowMax" + ax] - o["WindowMin" + ax]
        if is_relative:
            o["WindowMin" + ax] += p
            o["WindowMax" + ax] += p
        else:
            o["WindowMin" + ax] = p - w / 2.0
            o["WindowMax" + ax] = p + w / 2.0
    return g


def data_op(op, data, scalar, gpu_core, mask=None):
    raise AstraError("data_op (DATA_OPERATION_CUDA) is not part of this build")


def geom_postalignment(proj_geom, factor):
    raise AstraError("geom_postalignment is not part of this build")
