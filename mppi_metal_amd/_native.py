"""ctypes binding of the C ABI declared in include/mppi_hip.h.

The shared library is built in-tree by mppi_metal_amd/build.py.  Loading
fails loudly (ImportError) when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmppi_hip.so"

MPPI_HIP_OK = 0
STATUS_NAMES = {
    0: "OK", -1: "E_INVALID", -2: "E_NOT_INITIALIZED", -3: "E_PLUGIN",
    -4: "E_COMPILE", -5: "E_DEVICE", -6: "E_NO_MEMORY",
}
COV_DIAGONAL, COV_FULL = 0, 1
FP32, FP16, MIXED = 0, 1, 2
PATH_AUTO, PATH_TWO_KERNEL, PATH_FUSED = 0, 1, 2


class MppiHipConfig(ctypes.Structure):
    """Mirror of struct mppi_hip_config (include/mppi_hip.h)."""

    _fields_ = [
        ("state_dim", ctypes.c_uint32),
        ("control_dim", ctypes.c_uint32),
        ("horizon", ctypes.c_uint32),
        ("sample_count", ctypes.c_uint32),
        ("lambda_", ctypes.c_float),
        ("covariance_form", ctypes.c_uint32),
        ("sigma", ctypes.POINTER(ctypes.c_float)),
        ("u_min", ctypes.POINTER(ctypes.c_float)),
        ("u_max", ctypes.POINTER(ctypes.c_float)),
        ("numeric_mode", ctypes.c_uint32),
        ("device", ctypes.c_int32),
    ]


class MppiHipBatchTuning(ctypes.Structure):
    """Mirror of struct mppi_hip_batch_tuning (include/mppi_hip.h)."""

    _fields_ = [
        ("block", ctypes.c_uint32),
        ("persistent", ctypes.c_uint32),
        ("wgs_per_cu", ctypes.c_uint32),
        ("reduce_spt", ctypes.c_uint32),
        ("order", ctypes.c_uint32),
        ("launches", ctypes.c_uint32),
    ]


class MppiHipSurrogateConfig(ctypes.Structure):
    """Mirror of struct mppi_hip_surrogate_config (include/mppi_hip.h)."""

    _fields_ = [
        ("state_lo", ctypes.POINTER(ctypes.c_float)),
        ("state_hi", ctypes.POINTER(ctypes.c_float)),
        ("state_input", ctypes.POINTER(ctypes.c_uint32)),
        ("num_transitions", ctypes.c_uint32),
        ("batch", ctypes.c_uint32),
        ("steps", ctypes.c_uint32),
        ("learning_rate", ctypes.c_float),
        ("seed", ctypes.c_uint64),
    ]


SURROGATE_INPUT_BOX, SURROGATE_INPUT_PERIODIC, SURROGATE_INPUT_IGNORED = 0, 1, 2
SURROGATE_NORM_FLOATS = 80

class MppiHipRequest(ctypes.Structure):
    """Mirror of struct mppi_hip_request (include/mppi_hip.h)."""

    _fields_ = [
        ("arrival_s", ctypes.c_double),
        ("K", ctypes.c_uint32), ("H", ctypes.c_uint32),
        ("solves", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
        ("x0", ctypes.POINTER(ctypes.c_float)),
        ("u_nominal", ctypes.POINTER(ctypes.c_float)),
        ("cost_params", ctypes.c_void_p),
    ]


class MppiHipReplayStats(ctypes.Structure):
    """Mirror of struct mppi_hip_replay_stats (include/mppi_hip.h)."""

    _fields_ = [
        ("seconds", ctypes.c_double), ("sample_steps", ctypes.c_double),
        ("max_workspace_bytes", ctypes.c_double), ("batches", ctypes.c_uint64),
        ("max_batch", ctypes.c_uint32), ("admission_waits", ctypes.c_uint32),
    ]


_F = ctypes.POINTER(ctypes.c_float)
_U32P = ctypes.POINTER(ctypes.c_uint32)
_VP = ctypes.c_void_p
_CTX = ctypes.c_void_p

# (name, restype, argtypes) for every entry point of include/mppi_hip.h.
SIGNATURES = [
    ("mppi_hip_create", ctypes.c_int, [ctypes.POINTER(MppiHipConfig), ctypes.POINTER(_CTX)]),
    ("mppi_hip_destroy", None, [_CTX]),
    ("mppi_hip_last_error", ctypes.c_char_p, [_CTX]),
    ("mppi_hip_link_plugin", ctypes.c_int,
     [_CTX, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
      ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint64]),
    ("mppi_hip_link_plugin_source", ctypes.c_int,
     [_CTX, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
      ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint64]),
    ("mppi_hip_bind_params", ctypes.c_int, [_CTX, _VP, ctypes.c_size_t, _VP, ctypes.c_size_t]),
    ("mppi_hip_step", ctypes.c_int,
     [_CTX, ctypes.c_uint64, _F, _F, _VP, ctypes.c_size_t, _VP, ctypes.c_size_t, _F, _F]),
    ("mppi_hip_step_noalloc", ctypes.c_int, [_CTX, ctypes.c_uint64, _VP, _VP, _VP, _VP, _VP]),
    ("mppi_hip_simulate", ctypes.c_int,
     [_CTX, ctypes.c_uint32, ctypes.c_uint64, _F, _F, _VP, ctypes.c_size_t, _VP, ctypes.c_size_t,
      _F, _F, _F]),
    ("mppi_hip_simulate_batch", ctypes.c_int,
     [_CTX, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, _F, _F, _F, _F, _F, _F]),
    ("mppi_hip_read_costs", ctypes.c_int, [_CTX, ctypes.c_uint32, _F, ctypes.c_size_t]),
    ("mppi_hip_read_weights", ctypes.c_int, [_CTX, ctypes.c_uint32, _F, ctypes.c_size_t]),
    ("mppi_hip_set_reduce_samples_per_thread", ctypes.c_int, [_CTX, ctypes.c_uint32]),
    ("mppi_hip_set_step_path", ctypes.c_int, [_CTX, ctypes.c_uint32]),
    ("mppi_hip_set_wait_limit", ctypes.c_int, [_CTX, ctypes.c_uint32, ctypes.c_int]),
    ("mppi_hip_profile_enable", ctypes.c_int, [_CTX, ctypes.c_int]),
    ("mppi_hip_profile_read", ctypes.c_int,
     [_CTX, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
      ctypes.POINTER(ctypes.c_uint64)]),
    ("mppi_hip_batch_create", ctypes.c_int, [_CTX, ctypes.c_uint32, _U32P, _U32P, ctypes.POINTER(_VP)]),
    ("mppi_hip_batch_destroy", None, [_VP]),
    ("mppi_hip_batch_set_params", ctypes.c_int, [_VP, _VP, ctypes.c_size_t, _VP, ctypes.c_size_t]),
    ("mppi_hip_batch_set_tuning", ctypes.c_int, [_VP, ctypes.POINTER(MppiHipBatchTuning)]),
    ("mppi_hip_batch_get_tuning", ctypes.c_int, [_VP, ctypes.POINTER(MppiHipBatchTuning)]),
    ("mppi_hip_batch_autotune", ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    ("mppi_hip_batch_reset", ctypes.c_int, [_VP, ctypes.c_uint64, _F, _F]),
    ("mppi_hip_batch_run", ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.c_int, _F, _F, _F]),
    ("mppi_hip_batch_read", ctypes.c_int, [_VP, _F, _F, _F]),
    ("mppi_hip_surrogate_train", ctypes.c_int, [_CTX, ctypes.POINTER(MppiHipSurrogateConfig), _F]),
    ("mppi_hip_surrogate_evaluate", ctypes.c_int, [_CTX, ctypes.c_uint32, ctypes.c_uint64, _F, _F]),
    ("mppi_hip_surrogate_predict", ctypes.c_int, [_CTX, ctypes.c_uint32, _F, _F, _F]),
    ("mppi_hip_surrogate_use", ctypes.c_int, [_CTX, ctypes.c_int]),
    ("mppi_hip_surrogate_export", ctypes.c_int,
     [_CTX, _F, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), _F]),
    ("mppi_hip_surrogate_import", ctypes.c_int, [_CTX, _F, ctypes.c_size_t, _F]),
    ("mppi_hip_surrogate_gradients", ctypes.c_int, [_CTX, ctypes.c_uint32, _F, _F, _F, _F]),
    ("mppi_hip_surrogate_activation", ctypes.c_int, [_CTX, ctypes.c_uint32, _F, _F]),
    ("mppi_hip_server_create", ctypes.c_int, [_CTX, ctypes.c_uint32, ctypes.c_size_t, ctypes.POINTER(_VP)]),
    ("mppi_hip_server_destroy", None, [_VP]),
    ("mppi_hip_server_workspace_bytes", ctypes.c_int,
     [_VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_size_t)]),
    ("mppi_hip_server_admit", ctypes.c_int,
     [_VP, ctypes.c_uint32, ctypes.c_uint32, _F, _F, _VP, ctypes.c_size_t, _U32P]),
    ("mppi_hip_server_release", ctypes.c_int, [_VP, ctypes.c_uint32]),
    ("mppi_hip_server_solve", ctypes.c_int, [_VP, ctypes.c_uint32, _U32P, _F, _F]),
    ("mppi_hip_server_read", ctypes.c_int, [_VP, ctypes.c_uint32, _F, _F]),
    ("mppi_hip_server_usage", ctypes.c_int,
     [_VP, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t), _U32P]),
    ("mppi_hip_server_replay", ctypes.c_int,
     [_VP, ctypes.c_uint32, ctypes.POINTER(MppiHipRequest), ctypes.c_uint32,
      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(MppiHipReplayStats)]),
    ("mppi_hip_version", ctypes.c_char_p, []),
    ("mppi_hip_calibrate", ctypes.c_int,
     [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
      ctypes.POINTER(ctypes.c_double)]),
    ("mppi_hip_compile_check", ctypes.c_int,
     [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_char_p,
      ctypes.c_size_t]),
]

_lib = None


def load() -> ctypes.CDLL:
    """Load libmppi_hip.so (once) and declare the entry points."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m mppi_metal_amd.build`")
    lib = ctypes.CDLL(os.fspath(LIB_PATH))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
