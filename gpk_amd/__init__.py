"""gpk_amd — MI355X-native GPK per-step kinetics (HIP kernels behind a C ABI).

The compute path is libgpk_hip.so (gpk_amd/csrc); this package holds the
ctypes mirror of the reference GPKGrid/ThermoModel interface, the batched
instances (GPKBatch) and, in gpk_amd.adjoint, the velocity adjoint as a
torch.autograd.Function (imported on demand: it needs torch).
"""
from . import _abi
from ._abi import (aniso_ostwald, anisotropic_growth, dislocation, isotropic, model_defaults,
                   precipitation, recrystallisation, solidification, tensor_aniso, with_elastic)
from ._lib import GPKError, LIB_PATH
from .batch import GPKBatch
from .grid import GPKGrid, voronoi_host

__all__ = [
    "GPKGrid", "GPKBatch", "GPKError", "LIB_PATH", "voronoi_host", "isotropic", "anisotropic_growth",
    "solidification", "aniso_ostwald", "tensor_aniso", "recrystallisation", "dislocation",
    "precipitation", "with_elastic", "model_defaults", "_abi",
]
