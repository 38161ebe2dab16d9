"""MI355X-native ASTRA projection hot path, with the ``astra`` Python API.

The kept surface mirrors the reference's python/astra package for the GPU
projection path: data2d / data3d, create_projector / projector /
projector3d, algorithm (FP/BP/FBP/SIRT/CGLS 2D, FP3D/BP3D/FDK/SIRT3D/CGLS3D),
experimental direct FP/BP on DLPack tensors, and helpers.  All arithmetic
runs in libastra_amd (hand-written HIP kernels for gfx950).
"""
import os

from . import (algorithm, autograd, batched, composite, data2d, data3d, experimental, functions, log,
               projector, projector3d, pythonutils)
from ._core import set_gpu_index
from .creators import (astra_dict, create_backprojection, create_backprojection3d_gpu,
                       create_proj_geom, create_projector, create_reconstruction, create_sino,
                       create_sino3d_gpu, create_vol_geom)
from .functions import add_noise_to_sino, clear, data_op, geom_2vec, geom_size, move_vol_geom
from .log import AstraError
from .pythonutils import GPULink

try:  # scipy LinearOperator view of a projector (python/astra/__init__.py)
    from .optomo import OpTomo
except ImportError:  # pragma: no cover - scipy absent
    pass

__version__ = "2.5.99+mi355x"


def use_cuda():
    """True when a GPU and the native library are usable."""
    try:
        import torch
        from . import _lib
        _lib.load()
        return torch.cuda.is_available()
    except Exception:
        return False


def has_feature(name):
    return name in ("cuda", "gpu", "dlpack", "unchunked_3d", "hip")


def get_gpu_info(idx=-1):
    import torch
    i = 0 if idx < 0 else idx
    p = torch.cuda.get_device_properties(i)
    return "GPU #%d: %s, with %dMB" % (i, p.name, p.total_memory // (1024 * 1024))


if "ASTRA_GPU_INDEX" in os.environ:
    set_gpu_index([int(x) for x in os.environ["ASTRA_GPU_INDEX"].split(",")])
