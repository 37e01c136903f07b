"""ctypes front end of the CPU oracle — TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (the restatement in ``bf_oracle.hpp``) or
``oracle/_ref/libbfref.so`` (the reference's own CPU headers compiled in
place), which export the same C interface (``oracle_api.h``) under the
prefixes ``oracle_`` and ``ref_``.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s cpu_baseline leg import this module; the product package
``batchfit_amd`` never does.

Argument meaning follows the reference's Python entry point
(python/batchfit/__init__.py:63-112 and python/bindings.cpp:60-121): shared
x[M], fit-major y[N,M], optional p0[N,P] (None → data-only guess per row),
bounds [P]; outputs params[N,P], status int32[N], iterations uint32[N],
chi2[N] and, with uncertainties, covariance[N,P,P], stderr[N,P], uncertainty
status int32[N].
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import NamedTuple, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = {"gaussian": 0, "lorentzian": 1, "pseudo_voigt": 2, "two_lorentzian": 3,
          # test-only models of the restatement (liboracle.so), for closed-form checks
          "constant": 100, "exponential": 101, "affine": 102, "line": 103, "duplicate": 104}
PARAMS = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8,
          "constant": 1, "exponential": 1, "affine": 2, "line": 2, "duplicate": 2}


STOP_FIELDS = ("gradient", "gradient_tolerance", "step", "step_threshold", "reduction",
               "cost_threshold", "predicted", "predicted_threshold", "width_over_spacing",
               "lambda", "fired")


class Options(ctypes.Structure):
    _fields_ = [
        ("max_iterations", ctypes.c_uint32),
        ("defaults", ctypes.c_uint32),
        ("initial_lambda", ctypes.c_double),
        ("gradient_tolerance", ctypes.c_double),
        ("step_tolerance", ctypes.c_double),
        ("cost_tolerance", ctypes.c_double),
        ("compensated", ctypes.c_int32),
        ("scaled", ctypes.c_int32),
    ]


def make_options(max_iterations=100, initial_lambda=None, gradient_tolerance=None,
                 step_tolerance=None, cost_tolerance=None, compensated=False, scaled=False):
    given = (initial_lambda, gradient_tolerance, step_tolerance, cost_tolerance)
    defaults = sum(1 << i for i, v in enumerate(given) if v is None)
    values = [0.0 if v is None else float(v) for v in given]
    return Options(max_iterations, defaults, *values, int(compensated), int(scaled))


class Result(NamedTuple):
    params: np.ndarray
    status: np.ndarray
    iterations: np.ndarray
    chi2: np.ndarray
    covariance: Optional[np.ndarray] = None
    stderr: Optional[np.ndarray] = None
    uncertainty_status: Optional[np.ndarray] = None


def build(reference: bool = True) -> None:
    """Compile liboracle.so (and _ref/libbfref.so where the reference exists)."""
    target = "all" if reference else os.path.join(HERE, "liboracle.so")
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


_LIBS: dict = {}


def _ptr(array):
    return None if array is None else array.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One loaded oracle library. ``kind`` is "port" (the restatement) or "reference"."""

    def __init__(self, kind: str = "port"):
        if kind == "port":
            path, self.prefix = os.path.join(HERE, "liboracle.so"), "oracle_"
        elif kind == "reference":
            path, self.prefix = os.path.join(HERE, "_ref", "libbfref.so"), "ref_"
        else:
            raise ValueError(kind)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run `make -C oracle`")
        self.kind = kind
        self.lib = ctypes.CDLL(path)
        vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        for suffix in ("f32", "f64"):
            fn = getattr(self.lib, f"{self.prefix}fit_{suffix}")
            fn.argtypes = [i32, vp, sz, vp, sz, vp, vp, vp, ctypes.POINTER(Options), i32, i32,
                           vp, vp, vp, vp, vp, vp, vp]
            fn.restype = i32
            fn = getattr(self.lib, f"{self.prefix}guess_{suffix}")
            fn.argtypes = [i32, vp, sz, vp, sz, i32, vp, vp, vp]
            fn.restype = i32
            scalar = ctypes.c_float if suffix == "f32" else ctypes.c_double
            fn = getattr(self.lib, f"{self.prefix}model_{suffix}")
            fn.argtypes = [i32, scalar, vp, vp, vp]
            fn.restype = i32
            fn = getattr(self.lib, f"{self.prefix}cholesky_{suffix}")
            fn.argtypes = [vp, vp, i32, i32]
            fn.restype = i32
            fn = getattr(self.lib, f"{self.prefix}uncertainty_{suffix}")
            fn.argtypes = [i32, vp, sz, vp, ctypes.c_int32, scalar, vp, vp, vp, vp, vp]
            fn.restype = i32

    def fit_stops(self, x, y, model, p0=None, bounds=None, *, dtype=None, options=None,
                  threads=1):
        """The values of each row's last stop tests (restatement only), an
        [N, 11] float64 array with the columns of STOP_FIELDS."""
        if self.kind != "port":
            raise ValueError("stop diagnostics exist in the restatement only")
        dtype = np.dtype(dtype or np.result_type(np.asarray(x).dtype, np.asarray(y).dtype))
        x = np.ascontiguousarray(x, dtype=dtype)
        y = np.ascontiguousarray(y, dtype=dtype)
        if y.ndim == 1:
            y = y.reshape(1, -1)
        n, m = y.shape
        p = PARAMS[model]
        if p0 is not None:
            p0 = np.ascontiguousarray(np.broadcast_to(np.asarray(p0, dtype=dtype), (n, p)))
        lower, upper = _bounds(bounds, p, dtype)
        out = np.empty((n, 11), np.float64)
        opts = options if options is not None else make_options()
        fn = getattr(self.lib, f"oracle_fit_stops_{self._suffix(dtype)}")
        fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                       ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                       ctypes.POINTER(Options), ctypes.c_int, ctypes.c_void_p]
        fn.restype = ctypes.c_int
        if fn(MODELS[model], _ptr(x), m, _ptr(y), n, _ptr(p0), _ptr(lower), _ptr(upper),
              ctypes.byref(opts), int(threads), _ptr(out)) != 0:
            raise ValueError("stop diagnostics exist for the spectral models only")
        return out

    @staticmethod
    def _suffix(dtype):
        return "f32" if np.dtype(dtype) == np.float32 else "f64"

    def fit(self, x, y, model="lorentzian", p0=None, bounds=None, *, dtype=None,
            uncertainties=False, options=None, threads=1) -> Result:
        dtype = np.dtype(dtype or np.result_type(np.asarray(x).dtype, np.asarray(y).dtype))
        x = np.ascontiguousarray(x, dtype=dtype)
        y = np.ascontiguousarray(y, dtype=dtype)
        if y.ndim == 1:
            y = y.reshape(1, -1)
        n, m = y.shape
        p = PARAMS[model]
        if p0 is not None:
            p0 = np.ascontiguousarray(np.broadcast_to(np.asarray(p0, dtype=dtype), (n, p)))
        lower, upper = _bounds(bounds, p, dtype)
        params = np.empty((n, p), dtype)
        status = np.empty(n, np.int32)
        iterations = np.empty(n, np.uint32)
        chi2 = np.empty(n, dtype)
        cov = np.empty((n, p, p), dtype) if uncertainties else None
        se = np.empty((n, p), dtype) if uncertainties else None
        us = np.empty(n, np.int32) if uncertainties else None
        opts = options if options is not None else make_options()
        fn = getattr(self.lib, f"{self.prefix}fit_{self._suffix(dtype)}")
        rc = fn(MODELS[model], _ptr(x), m, _ptr(y), n, _ptr(p0), _ptr(lower), _ptr(upper),
                ctypes.byref(opts), int(uncertainties), int(threads), _ptr(params), _ptr(status),
                _ptr(iterations), _ptr(chi2), _ptr(cov), _ptr(se), _ptr(us))
        if rc != 0:
            raise ValueError("oracle rejected the model id")
        if uncertainties:
            return Result(params, status, iterations, chi2, cov, se, us)
        return Result(params, status, iterations, chi2)

    def guess(self, x, y, model, *, dtype=None, threads=1):
        dtype = np.dtype(dtype or np.result_type(np.asarray(x).dtype, np.asarray(y).dtype))
        x = np.ascontiguousarray(x, dtype=dtype)
        y = np.ascontiguousarray(y, dtype=dtype)
        if y.ndim == 1:
            y = y.reshape(1, -1)
        n, m = y.shape
        p = PARAMS[model]
        params = np.empty((n, p), dtype)
        baseline = np.empty(n, dtype)
        status = np.empty(n, np.int32)
        fn = getattr(self.lib, f"{self.prefix}guess_{self._suffix(dtype)}")
        fn(MODELS[model], _ptr(x), m, _ptr(y), n, int(threads), _ptr(params), _ptr(baseline),
           _ptr(status))
        return params, baseline, status

    def model(self, model, x, params, dtype=np.float64):
        dtype = np.dtype(dtype)
        p = np.ascontiguousarray(np.zeros(8, dtype))
        p[: len(params)] = params
        value = np.zeros(1, dtype)
        jac = np.zeros(8, dtype)
        getattr(self.lib, f"{self.prefix}model_{self._suffix(dtype)}")(
            MODELS[model], dtype.type(x), _ptr(p), _ptr(value), _ptr(jac))
        return value[0], jac[: PARAMS[model]]

    def cholesky(self, a, b, scaled=False, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype).copy()
        b = np.ascontiguousarray(b, dtype=dtype).copy()
        ok = getattr(self.lib, f"{self.prefix}cholesky_{self._suffix(dtype)}")(
            _ptr(a), _ptr(b), len(b), int(scaled))
        return bool(ok), a, b

    def uncertainty(self, model, x, params, status, chi2, bounds=None, dtype=np.float64):
        dtype = np.dtype(dtype)
        p = PARAMS[model]
        x = np.ascontiguousarray(x, dtype=dtype)
        params = np.ascontiguousarray(params, dtype=dtype)
        lower, upper = _bounds(bounds, p, dtype)
        cov = np.empty((p, p), dtype)
        se = np.empty(p, dtype)
        us = np.empty(1, np.int32)
        getattr(self.lib, f"{self.prefix}uncertainty_{self._suffix(dtype)}")(
            MODELS[model], _ptr(x), len(x), _ptr(params), int(status), dtype.type(chi2),
            _ptr(lower), _ptr(upper), _ptr(cov), _ptr(se), _ptr(us))
        return cov, se, int(us[0])


def _bounds(bounds, p, dtype):
    if bounds is None:
        return np.full(p, -np.inf, dtype), np.full(p, np.inf, dtype)
    lower, upper = bounds
    return (np.ascontiguousarray(np.broadcast_to(np.asarray(lower, dtype=dtype), (p,))),
            np.ascontiguousarray(np.broadcast_to(np.asarray(upper, dtype=dtype), (p,))))


def load(kind: str = "port") -> Oracle:
    if kind not in _LIBS:
        _LIBS[kind] = Oracle(kind)
    return _LIBS[kind]
