"""Toolbox-level functions (python/astra/astra.py, python/astra/astra_c.pyx):
credits, GPU selection and information, feature flags, and delete / info
of any object by id."""
from . import _core
from ._registry import algorithms, data2d, data3d, projectors2d, projectors3d

set_gpu_index = _core.set_gpu_index

# object kinds as the reference's managers name them (getType)
_KINDS = (("data2d", data2d), ("data3d", data3d), ("projector2d", projectors2d),
          ("projector3d", projectors3d), ("algorithm", algorithms))

# feature flags of hasFeature (src/Features.cpp:33-50) that hold for this build
_FEATURES = ("cuda", "projectors_scaled_as_line_integrals", "fan_cone_BP_density_weighting_by_default",
             "unpadded_GPULink")


def credits():
    """Print credits."""
    print("ASTRA Toolbox projection path for AMD Instinct MI355X (gfx950).\n"
          "The ASTRA Toolbox is developed at the University of Antwerp and CWI, Amsterdam;\n"
          "see https://www.astra-toolbox.com for its authors and references.")


def use_cuda():
    """True when the GPU path is usable (a GPU and the native library)."""
    try:
        import torch
        from . import _lib
        _lib.load()
        return bool(torch.cuda.is_available())
    except Exception:
        return False


def get_gpu_info(idx=-1):
    """Name and memory of GPU `idx` (-1: the current device)."""
    import torch
    i = torch.cuda.current_device() if idx < 0 else idx
    p = torch.cuda.get_device_properties(i)
    return "GPU #%d: %s, with %dMB" % (i, p.name, p.total_memory // (1024 * 1024))


def has_feature(feature):
    """A feature flag (src/Features.cpp): True for the flags this build
    honours, False for any other string."""
    return isinstance(feature, str) and feature in _FEATURES


def _ids(ids):
    if isinstance(ids, (str, bytes)) or not hasattr(ids, "__iter__"):
        return (ids,)
    return ids


def delete(ids):
    """Delete objects of any kind by id; unknown ids are ignored."""
    for i in _ids(ids):
        for _, reg in _KINDS:
            if i in reg.objects:
                reg.delete(i)
                break


def info(ids):
    """Print one line per existing object: its kind, id, initialised flag
    and description (python/astra/astra_c.pyx:137-149)."""
    for i in _ids(ids):
        for kind, reg in _KINDS:
            obj = reg.objects.get(i)
            if obj is not None:
                desc = obj.description() if hasattr(obj, "description") else type(obj).__name__
                print("%s\t%d \tv     %s" % (kind, i, desc))
                break
