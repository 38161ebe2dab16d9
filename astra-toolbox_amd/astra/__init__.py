"""MI355X-native ASTRA projection hot path, with the ``astra`` Python API.

The kept surface mirrors the reference's python/astra package for the GPU
projection path: data2d / data3d, create_projector / projector /
projector3d, algorithm (FP/BP/FBP/SIRT/CGLS 2D, FP3D/BP3D/FDK/SIRT3D/CGLS3D),
experimental direct FP/BP on DLPack tensors, and helpers.  All arithmetic
runs in libastra_amd (hand-written HIP kernels for gfx950).
"""
import os

from . import (algorithm, autograd, batched, composite, data2d, data3d, experimental, functions, log,
               matrix, plugin, plugins, projector, projector3d, pythonutils)
from . import astra
from . import matlab
from . import matlab as m
from .astra import get_gpu_info, has_feature, set_gpu_index, use_cuda
from .extrautils import clipCircle
from .tests import test, test_CUDA, test_noCUDA
from .creators import (astra_dict, create_backprojection, create_backprojection3d_gpu,
                       create_proj_geom, create_projector, create_reconstruction, create_sino,
                       create_sino3d_gpu, create_vol_geom)
from .functions import (add_noise_to_sino, clear, data_op, geom_2vec, geom_postalignment, geom_size,
                        move_vol_geom)
from .log import AstraError
from .pythonutils import GPULink

try:  # scipy LinearOperator view of a projector (python/astra/__init__.py)
    from .optomo import OpTomo
except ImportError:  # pragma: no cover - scipy absent
    pass

__version__ = "2.5.99+mi355x"


if "ASTRA_GPU_INDEX" in os.environ:
    set_gpu_index([int(x) for x in os.environ["ASTRA_GPU_INDEX"].split(",")])
