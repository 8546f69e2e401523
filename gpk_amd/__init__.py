"""gpk_amd — MI355X-native GPK per-step kinetics (HIP kernels behind a C ABI).

The compute path is libgpk_hip.so (gpk_amd/csrc); this package holds the
ctypes mirror of the reference GPKGrid/ThermoModel interface.
"""
from . import _abi
from ._abi import (aniso_ostwald, anisotropic_growth, isotropic, model_defaults,
                   solidification, tensor_aniso)
from ._lib import GPKError, LIB_PATH
from .grid import GPKGrid, voronoi_host

__all__ = [
    "GPKGrid", "GPKError", "LIB_PATH", "voronoi_host", "isotropic", "anisotropic_growth",
    "solidification", "aniso_ostwald", "tensor_aniso", "model_defaults", "_abi",
]
