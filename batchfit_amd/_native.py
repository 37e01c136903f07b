"""ctypes binding of libbatchfit_hip.so (include/batchfit_hip.h).

The shared library is built in-tree by ``batchfit_amd.build`` (hipcc, gfx950).
Loading fails loudly when it is missing: there is no CPU fallback behind the
HIP device path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libbatchfit_hip.so")

MODELS = {"gaussian": 0, "lorentzian": 1, "pseudo_voigt": 2, "two_lorentzian": 3}
PARAMETERS = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8}

BF_OK, BF_INVALID_ARGUMENT, BF_RUNTIME_ERROR, BF_UNAVAILABLE, BF_TIMEOUT = 0, 1, 2, 3, 4
COMPENSATED, SCALED_CHOLESKY, MIXED_FP64 = 1, 2, 4
MAP_AUTO, MAP_LANE, MAP_GROUP, MAP_LATENCY = 0, 1, 2, 3

# Every symbol include/batchfit_hip.h declares.
EXPORTS = (
    "bf_hip_abi_version", "bf_hip_available", "bf_hip_default_config", "bf_hip_fit",
    "bf_hip_fit_device", "bf_hip_fit_device_noalloc", "bf_hip_guess", "bf_hip_guess_form",
    "bf_hip_model_eval",
    "bf_hip_plan", "bf_hip_plan_batch", "bf_hip_set_timing",
    "bf_hip_last_timing", "bf_hip_malloc", "bf_hip_free", "bf_hip_memcpy", "bf_hip_synchronize",
    "bf_hip_ragged_create", "bf_hip_ragged_run", "bf_hip_ragged_timing", "bf_hip_ragged_destroy",
    "bf_hip_synthesize", "bf_serve_create", "bf_serve_submit", "bf_serve_wait", "bf_serve_stats",
    "bf_serve_destroy", "bf_npy_inspect", "bf_npy_chunk_fits", "bf_hip_fit_npy",
)


class Config(ctypes.Structure):
    """metal::detail::Config layout (include/batchfit/metal/types.hpp:7-15), 108 bytes."""

    _fields_ = [
        ("fits", ctypes.c_uint32),
        ("points", ctypes.c_uint32),
        ("max_iterations", ctypes.c_uint32),
        ("initial_lambda", ctypes.c_float),
        ("gradient_tolerance", ctypes.c_float),
        ("step_tolerance", ctypes.c_float),
        ("cost_tolerance", ctypes.c_float),
        ("lower", ctypes.c_float * 8),
        ("upper", ctypes.c_float * 8),
        ("numerics", ctypes.c_uint32),
        ("minimum_spacing", ctypes.c_float),
        ("varied_x", ctypes.c_uint32),
        ("width_factor", ctypes.c_float),
    ]


assert ctypes.sizeof(Config) == 108


class Launch(ctypes.Structure):
    _fields_ = [
        ("mapping", ctypes.c_uint32),
        ("group_lanes", ctypes.c_uint32),
        ("point_major", ctypes.c_uint32),
        ("streamed", ctypes.c_uint32),
        ("stream", ctypes.c_void_p),
    ]


class HipError(RuntimeError):
    """A HIP runtime failure reported by the library (message names the call)."""


_lib = None


def library() -> ctypes.CDLL:
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY):
        raise ImportError(
            f"{LIBRARY} is missing: build it with `python -m batchfit_amd.build` "
            "(hipcc --offload-arch=gfx950); the HIP path has no CPU fallback")
    lib = ctypes.CDLL(LIBRARY)
    vp, u32, i32, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t
    cfgp, lp = ctypes.POINTER(Config), ctypes.POINTER(Launch)
    lib.bf_hip_abi_version.restype = i32
    lib.bf_hip_available.restype = i32
    lib.bf_hip_default_config.argtypes = [cfgp]
    lib.bf_hip_default_config.restype = None
    fit_args = [u32, cfgp, lp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_char_p, sz]
    lib.bf_hip_fit.argtypes = fit_args
    lib.bf_hip_fit.restype = i32
    lib.bf_hip_fit_device.argtypes = fit_args
    lib.bf_hip_fit_device.restype = i32
    lib.bf_hip_fit_device_noalloc.argtypes = fit_args
    lib.bf_hip_fit_device_noalloc.restype = i32
    lib.bf_hip_guess.argtypes = [u32, u32, u32, vp, vp, vp, vp, vp, ctypes.c_char_p, sz]
    lib.bf_hip_guess.restype = i32
    lib.bf_hip_guess_form.argtypes = [u32, u32, u32, u32, vp, vp, vp, vp, vp, ctypes.c_char_p, sz]
    lib.bf_hip_guess_form.restype = i32
    lib.bf_hip_model_eval.argtypes = [u32, u32, vp, vp, vp, vp, ctypes.c_char_p, sz]
    lib.bf_hip_model_eval.restype = i32
    lib.bf_hip_plan.argtypes = [u32, u32, u32, vp, ctypes.POINTER(u32), ctypes.POINTER(u32),
                                ctypes.c_char_p, sz]
    lib.bf_hip_plan.restype = i32
    lib.bf_hip_plan_batch.argtypes = [u32, u32, u32, u32, vp, ctypes.POINTER(u32),
                                      ctypes.POINTER(u32), ctypes.c_char_p, sz]
    lib.bf_hip_plan_batch.restype = i32
    lib.bf_hip_set_timing.argtypes = [i32]
    lib.bf_hip_set_timing.restype = i32
    lib.bf_hip_last_timing.argtypes = [ctypes.POINTER(ctypes.c_double), i32]
    lib.bf_hip_last_timing.restype = i32
    lib.bf_hip_malloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), sz, ctypes.c_char_p, sz]
    lib.bf_hip_malloc.restype = i32
    lib.bf_hip_free.argtypes = [vp]
    lib.bf_hip_free.restype = i32
    lib.bf_hip_memcpy.argtypes = [vp, vp, sz, i32, ctypes.c_char_p, sz]
    lib.bf_hip_memcpy.restype = i32
    lib.bf_hip_synchronize.argtypes = [ctypes.c_char_p, sz]
    lib.bf_hip_synchronize.restype = i32
    _lib = lib
    return lib


def check(rc: int, buffer) -> None:
    """Map a library return code to the reference's exception types."""
    if rc == BF_OK:
        return
    message = buffer.value.decode(errors="replace") if buffer is not None else ""
    if rc == BF_INVALID_ARGUMENT:
        raise ValueError(message)
    if rc == BF_UNAVAILABLE:
        raise RuntimeError(message)
    raise HipError(message)


def error_buffer():
    return ctypes.create_string_buffer(512)


def default_config() -> Config:
    cfg = Config()
    library().bf_hip_default_config(ctypes.byref(cfg))
    return cfg
