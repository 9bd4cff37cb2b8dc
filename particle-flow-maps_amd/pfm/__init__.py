"""pfm -- MI355X-native particle-flow-map solver step (HIP kernels behind a C ABI).

Reference-compatible call surfaces live in ``pfm.d3`` (modules ``run_3D``,
``init_conditions``, ``taichi_utils``, ``mgpcg``, ``hyperparameters`` mirroring
``3D/*.py``) and ``pfm.d2`` (mirroring ``2D/*.py``).  ``pfm.simulation`` holds
the step drivers that follow the reference main loops.
"""
from . import _lib
from ._lib import PfmError
from .core import face_shape, default_dx
from .mgpcg import MGPCG, MGPCG_2, MGPCG_3

__all__ = ["MGPCG", "MGPCG_2", "MGPCG_3", "PfmError", "face_shape", "default_dx", "load_library"]


def load_library(require_gpu: bool = True):
    """Open libpfm.so (raises when it is missing, or no GPU when required)."""
    return _lib.load(require_gpu=require_gpu)
