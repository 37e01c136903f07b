"""Batched bounded Levenberg–Marquardt fitting of Gaussian, Lorentzian,
pseudo-Voigt and paired-Lorentzian spectra on AMD Instinct MI355X (gfx950).

``fit`` keeps the argument meaning, validation and result types of the
reference's ``batchfit.fit`` (python/batchfit/__init__.py:63-112) and adds the
device ``"hip"``, which runs hand-written CDNA4 kernels through the C ABI of
``libbatchfit_hip.so`` (include/batchfit_hip.h). Like the reference's explicit
GPU device, ``"hip"`` computes in float32 and returns float32 arrays
(python/README.md:56-59). The reference's CPU device is not part of this
package: ``device="cpu"`` is rejected with a message rather than served by a
fallback, and so is ``device="auto"`` on float64 or integer inputs, which the
reference's "auto" keeps on that CPU path (python/batchfit/__init__.py:84-86).
"""
from __future__ import annotations

import ctypes
from enum import IntEnum
from typing import NamedTuple

import numpy as np

from . import _native

__version__ = "0.1.0"
__all__ = ["fit", "FitResult", "FitResultWithUncertainties", "FitStatus", "UncertaintyStatus",
           "hip_available", "NUMERICS_DEFAULT", "__version__"]

_PARAMETERS = dict(_native.PARAMETERS)

# Compensated FP32 partials + FP64 combine/solve/accept/covariance.
NUMERICS_DEFAULT = _native.COMPENSATED | _native.MIXED_FP64


class FitStatus(IntEnum):
    """cpu/solver.hpp:16-23."""
    converged = 0
    max_iterations = 1
    singular_matrix = 2
    invalid_input = 3
    degenerate_data = 4
    unresolved_peak = 5


class UncertaintyStatus(IntEnum):
    """uncertainty_types.hpp:3-10."""
    available = 0
    not_converged = 1
    insufficient_dof = 2
    singular_matrix = 3
    active_bound = 4
    invalid_input = 5


class FitResult(NamedTuple):
    params: np.ndarray
    status: np.ndarray
    iterations: np.ndarray
    chi2: np.ndarray


class FitResultWithUncertainties(NamedTuple):
    params: np.ndarray
    status: np.ndarray
    iterations: np.ndarray
    chi2: np.ndarray
    covariance: np.ndarray
    stderr: np.ndarray
    uncertainty_status: np.ndarray


def hip_available() -> bool:
    """Whether the library is built and a gfx950 device is usable."""
    try:
        return bool(_native.library().bf_hip_available())
    except ImportError:
        return False


def _real_array(value, name):
    array = np.asarray(value)
    if array.dtype.kind not in "fiu":
        raise TypeError(f"{name} must contain real numeric values")
    return array


def _validate_batch(x, y, model, p0, bounds):
    """Argument checks of the reference's ``fit`` (python/batchfit/__init__.py:63-112,
    python/bindings.cpp:153-161); returns float32 x[M], y[N,M], p0[N,P] or None
    and the broadcast bounds lower[P], upper[P]."""
    if model not in _PARAMETERS:
        raise ValueError(f"unknown model {model!r}; choose one of {tuple(_PARAMETERS)}")
    x, y = _real_array(x, "x"), _real_array(y, "y")
    if x.ndim != 1 or not x.size:
        raise ValueError("x must have nonempty shape [M]")
    if y.ndim == 1:
        y = y.reshape(1, -1)
    if y.ndim != 2 or y.shape[1] != x.size:
        raise ValueError("y must have shape [N,M] matching x[M]")
    dtype = np.dtype(np.float32)
    x = np.require(x, dtype=dtype, requirements=["C", "A"])
    y = np.require(y, dtype=dtype, requirements=["C", "A"])
    n, p = y.shape[0], _PARAMETERS[model]
    if n >= 2**32:
        raise ValueError("at most 2**32-1 fits per call")
    if p0 is not None:
        p0 = _real_array(p0, "p0")
        if p0.shape == (p,):
            p0 = np.broadcast_to(p0, (n, p))
        if p0.shape != (n, p):
            raise ValueError(f"p0 must have shape [{p}] or [{n},{p}]")
        p0 = np.require(p0, dtype=dtype, requirements=["C", "A"])
    if bounds is None:
        lower, upper = np.full(p, -np.inf, dtype=dtype), np.full(p, np.inf, dtype=dtype)
    else:
        if not isinstance(bounds, (tuple, list)) or len(bounds) != 2:
            raise ValueError("bounds must be (lower, upper), each scalar or [P]")
        lower, upper = (_real_array(value, "bounds") for value in bounds)
        if lower.shape not in ((), (p,)) or upper.shape not in ((), (p,)):
            raise ValueError(f"each bound must be scalar or have shape [{p}]")
        maximum = np.finfo(np.float32).max
        if any(np.any(np.isfinite(bound) & (np.abs(bound) > maximum)) for bound in (lower, upper)):
            raise ValueError("finite bounds must be representable in the fitting dtype")
        lower = np.require(np.broadcast_to(lower, (p,)), dtype=dtype, requirements=["C", "A"])
        upper = np.require(np.broadcast_to(upper, (p,)), dtype=dtype, requirements=["C", "A"])
    # python/bindings.cpp:153-161: invalid bounds are an argument error.
    if (np.any(np.isnan(lower)) or np.any(np.isnan(upper)) or np.any(lower > upper)
            or np.any(lower == np.inf) or np.any(upper == -np.inf)):
        raise ValueError("invalid parameter bounds")
    return x, y, p0, lower, upper


def _make_config(n, m, lower, upper, max_iterations, numerics):
    cfg = _native.default_config()
    cfg.fits, cfg.points = int(n), int(m)
    cfg.max_iterations = int(max_iterations)
    cfg.numerics = int(numerics)
    for k in range(len(lower)):
        cfg.lower[k], cfg.upper[k] = float(lower[k]), float(upper[k])
    return cfg


_MAPPINGS = {"auto": (0, 0, 0), "lane": (1, 0, 0), "group8": (2, 8, 0), "group16": (2, 16, 0),
             "group32": (2, 32, 0), "group64": (2, 64, 0), "wave": (2, 64, 0),
             "group4_streamed": (2, 4, 1), "group8_streamed": (2, 8, 1), "group16_streamed": (2, 16, 1),
             "group32_streamed": (2, 32, 1), "group64_streamed": (2, 64, 1),
             "group128": (2, 128, 0), "group256": (2, 256, 0), "workgroup": (2, 256, 0),
             "group128_streamed": (2, 128, 1), "group256_streamed": (2, 256, 1),
             "latency": (3, 0, 0)}


def plan(model, points, *, numerics=None, mapping="auto", fits=0):
    """The mapping fit() uses for ``points`` per spectrum: ``(lanes, tile)``,
    lanes per fit (1 = one lane per fit) and points per lane held in
    registers (0 = read from memory on every pass). ``fits`` is the batch
    size, which only mapping="latency" depends on. Host-only: no device is
    needed and nothing is launched."""
    if model not in _PARAMETERS:
        raise ValueError(f"unknown model {model!r}; choose one of {tuple(_PARAMETERS)}")
    if mapping not in _MAPPINGS:
        raise ValueError(f"mapping must be one of {tuple(_MAPPINGS)}")
    lib = _native.library()
    launch = _native.Launch()
    launch.mapping, launch.group_lanes, launch.streamed = _MAPPINGS[mapping]
    lanes, tile = ctypes.c_uint32(), ctypes.c_uint32()
    err = _native.error_buffer()
    _native.check(lib.bf_hip_plan_batch(_native.MODELS[model], int(points), int(fits),
                                        NUMERICS_DEFAULT if numerics is None else int(numerics),
                                        ctypes.byref(launch), ctypes.byref(lanes),
                                        ctypes.byref(tile), err, len(err)), err)
    return lanes.value, tile.value


def fit(x, y, model="lorentzian", p0=None, bounds=None, device="auto", *, uncertainties=False,
        numerics=None, mapping="auto", max_iterations=100):
    """Fit shared x[M] and y[N,M]; return params[N,P], status/iterations/chi2[N].

    A single y[M] is accepted as N=1. ``device`` is "auto" or "hip"; both run
    on the GPU in float32 and never fall back to a CPU path. As in the
    reference, "auto" takes float32 batches only (the reference keeps float64
    on its CPU path, so "auto" raises ValueError for float64 or integer
    inputs here); "hip" converts any real input to float32, as the
    reference's explicit GPU device does (python/README.md:56-59). ``p0=None`` runs
    the data-only initial guess on the device. ``uncertainties=True`` adds
    covariance[N,P,P], stderr[N,P] and uncertainty_status[N].

    Keyword-only extensions of the HIP device: ``numerics`` (bit0 compensated
    partials, bit1 scaled Cholesky, bit2 FP64 combine/solve; default 5),
    ``mapping`` ("auto", "lane", "group8", "group16", "group32", "group64", and
    "group128" / "group256" ("workgroup"), a workgroup of 2 or 4 waves per fit;
    a "_streamed" suffix reads the points from memory on every pass instead of
    registers, with the same results; "group4_streamed" streams 4-lane groups;
    "latency" gives a batch too small to fill the device 64, 128 or 256 lanes
    per fit, for a shorter call, and is "auto" otherwise) and
    ``max_iterations`` (default 100, cpu/solver.hpp:27).
    """
    if model not in _PARAMETERS:
        raise ValueError(f"unknown model {model!r}; choose one of {tuple(_PARAMETERS)}")
    if device == "cpu":
        raise ValueError("device 'cpu' is the reference CPU path, which this package does not "
                         "include; use device='hip'")
    if device not in ("auto", "hip"):
        raise ValueError("device must be 'auto' or 'hip'")
    if not isinstance(uncertainties, (bool, np.bool_)):
        raise TypeError("uncertainties must be a boolean")
    if mapping not in _MAPPINGS:
        raise ValueError(f"mapping must be one of {tuple(_MAPPINGS)}")
    if device == "auto":
        # python/batchfit/__init__.py:84-86: "auto" sends a float32 batch to
        # the GPU and keeps any other dtype (float64, integers) on the float64
        # CPU path, which this package does not include. Converting it here
        # would change the results' precision and dtype without a word.
        common = np.result_type(_real_array(x, "x").dtype, _real_array(y, "y").dtype)
        if common != np.dtype(np.float32):
            raise ValueError(
                f"device='auto' fits {common} inputs on the float64 CPU path of the reference, "
                "which this package does not include; pass float32 arrays, or device='hip' to "
                "fit in float32 (inputs converted, float32 results)")
    x, y, p0, lower, upper = _validate_batch(x, y, model, p0, bounds)
    dtype = np.dtype(np.float32)
    n, p = y.shape[0], _PARAMETERS[model]

    lib = _native.library()
    if not lib.bf_hip_available():
        raise RuntimeError("HIP fitting is unavailable in this build or on this device")
    cfg = _make_config(n, x.size, lower, upper, max_iterations,
                       NUMERICS_DEFAULT if numerics is None else numerics)
    launch = _native.Launch()
    launch.mapping, launch.group_lanes, launch.streamed = _MAPPINGS[mapping]

    params = np.empty((n, p), dtype)
    status = np.empty(n, np.int32)
    iterations = np.empty(n, np.uint32)
    chi2 = np.empty(n, dtype)
    cov = se = ustatus = None
    if uncertainties:
        cov = np.empty((n, p, p), dtype)
        se = np.empty((n, p), dtype)
        ustatus = np.empty(n, np.int32)

    def ptr(a):
        return None if a is None else a.ctypes.data_as(ctypes.c_void_p)

    err = _native.error_buffer()
    rc = lib.bf_hip_fit(_native.MODELS[model], ctypes.byref(cfg), ctypes.byref(launch), ptr(x),
                        ptr(y), ptr(p0), ptr(params), ptr(status), ptr(iterations), ptr(chi2),
                        ptr(cov), ptr(se), ptr(ustatus), err, len(err))
    _native.check(rc, err)
    if uncertainties:
        return FitResultWithUncertainties(params, status, iterations, chi2, cov, se, ustatus)
    return FitResult(params, status, iterations, chi2)
