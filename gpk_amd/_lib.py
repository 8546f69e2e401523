"""Loader for libgpk_hip.so (the HIP library built in-tree next to this file).

There is no CPU fallback: if the library is missing or fails to load, every
entry point raises.  Build it with ``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C gpk_amd/csrc``.
"""
import ctypes as C
import os

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpk_hip.so")

_lib = None


class GPKError(RuntimeError):
    """Raised for a non-zero status from the C ABI (message from gpk_last_error)."""

    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GPKError(_abi.GPK_E_NO_DEVICE,
                       "libgpk_hip.so is not built (expected at %s)" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    H = C.c_void_p
    L.gpk_last_error.restype = C.c_char_p
    L.gpk_version.restype = C.c_char_p
    L.gpk_model_defaults.argtypes = [C.c_int32, C.POINTER(_abi.ModelDesc)]
    L.gpk_model_defaults.restype = None
    L.gpk_create.argtypes = [C.POINTER(_abi.GridDesc), C.POINTER(_abi.ModelDesc), C.POINTER(H)]
    L.gpk_destroy.argtypes = [H]
    for name in ("gpk_upload", "gpk_upload_device"):
        getattr(L, name).argtypes = [H, C.c_int, vp, C.c_size_t]
    for name in ("gpk_download", "gpk_download_device"):
        getattr(L, name).argtypes = [H, C.c_int, vp, C.c_size_t]
    L.gpk_step.argtypes = [H, C.c_float, C.c_float, C.POINTER(C.c_double),
                           C.POINTER(_abi.StepStats)]
    L.gpk_step_many.argtypes = [H, C.c_int, C.c_float, C.c_float, C.POINTER(_abi.StepStats)]
    L.gpk_run.argtypes = [H, C.c_double, C.c_float, C.POINTER(C.c_int)]
    L.gpk_set_model.argtypes = [H, C.POINTER(_abi.ModelDesc)]
    L.gpk_get_time.argtypes = [H, C.POINTER(C.c_double)]
    L.gpk_set_time.argtypes = [H, C.c_double]
    L.gpk_max_dt.argtypes = [H, C.POINTER(C.c_double)]
    L.gpk_phase_volume.argtypes = [H, C.c_int32, C.POINTER(C.c_double)]
    L.gpk_n_grains.argtypes = [H, C.POINTER(C.c_int)]
    L.gpk_mean_grain_radius.argtypes = [H, C.POINTER(C.c_double)]
    L.gpk_solid_fraction.argtypes = [H, C.POINTER(C.c_float)]
    L.gpk_interface_fraction.argtypes = [H, C.POINTER(C.c_float)]
    L.gpk_solute_sum.argtypes = [H, C.POINTER(C.c_double)]
    L.gpk_device_bytes.argtypes = [H, C.POINTER(C.c_size_t)]
    L.gpk_batch_create.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_float,
                                   C.c_float, C.c_float, C.c_int, C.c_int, C.POINTER(_abi.ModelDesc),
                                   C.POINTER(H)]
    L.gpk_batch_destroy.argtypes = [H]
    L.gpk_batch_upload.argtypes = [H, C.c_int, vp, C.c_size_t]
    L.gpk_batch_download.argtypes = [H, C.c_int, vp, C.c_size_t]
    L.gpk_batch_step.argtypes = [H, C.c_int, C.c_float, C.c_float]
    L.gpk_batch_stats_get.argtypes = [H, C.POINTER(_abi.BatchStats)]
    L.gpk_batch_instances.argtypes = [H, vp, vp, vp]
    L.gpk_batch_autotune.argtypes = [H, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.gpk_batch_set_launch.argtypes = [H, C.c_int, C.c_int]
    L.gpk_velocity_field.argtypes = [H, C.POINTER(C.c_double), vp]
    L.gpk_velocity_adjoint.argtypes = [H, C.POINTER(C.c_double), vp, C.POINTER(C.c_double)]
    L.gpk_velocity_adjoint_noalloc.argtypes = [H, C.POINTER(C.c_double), vp, C.POINTER(C.c_double)]
    L.gpk_topology_simple_move.argtypes = [H, C.c_int, C.c_int32, C.POINTER(C.c_int)]
    L.gpk_voronoi_init.argtypes = [H, C.c_int, C.c_uint32]
    L.gpk_voronoi_host.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32, vp, vp]
    L.gpk_sync.argtypes = [H]
    L.gpk_set_profiling.argtypes = [H, C.c_int]
    L.gpk_kernel_timing.argtypes = [H, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_double)]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().gpk_last_error().decode(errors="replace")
        raise GPKError(rc, msg or "gpk call failed")
    return rc


# every symbol include/gpk_c.h declares
EXPORTS = (
    "gpk_last_error", "gpk_version", "gpk_model_defaults", "gpk_create", "gpk_destroy",
    "gpk_upload", "gpk_download", "gpk_upload_device", "gpk_download_device", "gpk_step",
    "gpk_step_many", "gpk_run", "gpk_set_model", "gpk_get_time", "gpk_set_time", "gpk_max_dt",
    "gpk_phase_volume", "gpk_n_grains", "gpk_mean_grain_radius", "gpk_solid_fraction",
    "gpk_interface_fraction", "gpk_solute_sum", "gpk_device_bytes", "gpk_topology_simple_move",
    "gpk_voronoi_init", "gpk_voronoi_host", "gpk_sync", "gpk_set_profiling", "gpk_kernel_timing",
    "gpk_velocity_field", "gpk_velocity_adjoint", "gpk_velocity_adjoint_noalloc",
    "gpk_batch_create", "gpk_batch_destroy", "gpk_batch_upload", "gpk_batch_download", "gpk_batch_step",
    "gpk_batch_stats_get", "gpk_batch_instances", "gpk_batch_autotune", "gpk_batch_set_launch",
)
