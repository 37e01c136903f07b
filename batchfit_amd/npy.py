"""NPY inputs and the streaming fit pipeline (``bf_npy_*``, ``bf_hip_fit_npy``).

``fit_npy(x_path, y_path, model, p0_path=None, bounds=None, ...)`` fits every
row of an NPY file y[N,M] (float32 or float64) against x[M] without loading
the file: chunks are read by several threads into a ring of pinned slots,
fitted on the device (float64 is narrowed there) and returned in order — the
pipeline of the reference's ``metal::fit_npy`` (src/metal/pipeline.cpp:70-265)
with the guess and the conversion moved to the device. Without a ``consumer``
the results are gathered into one FitResult; with one, ``consumer(first_fit,
chunk_result)`` is called for every chunk in order.
"""
from __future__ import annotations

import ctypes
import os
from typing import Callable, NamedTuple, Optional

import numpy as np

from . import _native

_DTYPES = {0: np.dtype("<f4"), 1: np.dtype("<f8"), 2: np.dtype("<u4")}


class NpyHeader(NamedTuple):
    shape: tuple
    dtype: np.dtype
    offset: int


class NpyOptions(ctypes.Structure):
    _fields_ = [("chunk_fits", ctypes.c_uint64), ("in_flight", ctypes.c_uint32),
                ("uncertainties", ctypes.c_int32), ("budget_bytes", ctypes.c_uint64)]


class NpyStats(ctypes.Structure):
    _fields_ = [("fits", ctypes.c_uint64), ("chunks", ctypes.c_uint64),
                ("chunk_fits", ctypes.c_uint64), ("max_in_flight", ctypes.c_uint64),
                ("max_mapped_bytes", ctypes.c_uint64), ("shared_payload_bytes", ctypes.c_uint64),
                ("wall_seconds", ctypes.c_double), ("prepare_seconds", ctypes.c_double),
                ("device_seconds", ctypes.c_double), ("readback_seconds", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_CONSUMER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)


def _setup(lib):
    if getattr(lib, "_npy_ready", False):
        return
    vp, u32, u64, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t
    lib.bf_npy_inspect.argtypes = [ctypes.c_char_p, ctypes.POINTER(u64), ctypes.POINTER(u32),
                                   ctypes.POINTER(u32), ctypes.POINTER(u64), ctypes.c_char_p, sz]
    lib.bf_npy_inspect.restype = ctypes.c_int
    lib.bf_npy_chunk_fits.argtypes = [u64, u64, u32, u32, u32, ctypes.POINTER(NpyOptions)]
    lib.bf_npy_chunk_fits.restype = u64
    lib.bf_hip_fit_npy.argtypes = [u32, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                   ctypes.POINTER(_native.Config), ctypes.POINTER(NpyOptions),
                                   _CONSUMER, vp, ctypes.POINTER(NpyStats), ctypes.c_char_p, sz]
    lib.bf_hip_fit_npy.restype = ctypes.c_int
    lib._npy_ready = True


def inspect(path) -> NpyHeader:
    """Shape, dtype and payload offset of an NPY v1.0 file (npy::inspect)."""
    lib = _native.library()
    _setup(lib)
    shape = (ctypes.c_uint64 * 2)()
    rank, dtype, offset = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
    err = _native.error_buffer()
    _native.check(lib.bf_npy_inspect(os.fsencode(path), shape, ctypes.byref(rank),
                                     ctypes.byref(dtype), ctypes.byref(offset), err, len(err)), err)
    return NpyHeader(tuple(int(shape[k]) for k in range(rank.value)), _DTYPES[dtype.value],
                     int(offset.value))


def chunk_fits(fits, points, parameters, input_bytes=4, p0_bytes=0, *, in_flight=3,
               uncertainties=False, chunk=0, budget_bytes=0) -> int:
    lib = _native.library()
    _setup(lib)
    opt = NpyOptions(int(chunk), int(in_flight), int(bool(uncertainties)), int(budget_bytes))
    return int(lib.bf_npy_chunk_fits(int(fits), int(points), int(parameters), int(input_bytes),
                                     int(p0_bytes), ctypes.byref(opt)))


def fit_npy(x_path, y_path, model="lorentzian", p0_path=None, bounds=None, *,
            uncertainties=False, chunk_fits=0, in_flight=3, budget_bytes=0, max_iterations=100,
            numerics=None, consumer: Optional[Callable] = None):
    """Fit y[N,M] from `y_path` against x[M] from `x_path`; returns
    ``(result, stats)`` where result is a FitResult (or
    FitResultWithUncertainties) for all N rows, or None when a consumer took
    the chunks, and stats is a dict of the pipeline counters."""
    from . import (FitResult, FitResultWithUncertainties, NUMERICS_DEFAULT, _PARAMETERS,
                   _make_config, _validate_batch)
    if model not in _PARAMETERS:
        raise ValueError(f"unknown model {model!r}; choose one of {tuple(_PARAMETERS)}")
    if not 1 <= int(in_flight) <= 8:
        raise ValueError("in_flight must be 1..8")
    lib = _native.library()
    _setup(lib)
    yh = inspect(y_path)
    if len(yh.shape) != 2:
        raise ValueError("y must have shape [N,M]")
    n, m = yh.shape
    p = _PARAMETERS[model]
    _, _, _, lower, upper = _validate_batch(np.zeros(m, np.float32), np.zeros((0, m), np.float32),
                                            model, None, bounds)
    cfg = _make_config(n, m, lower, upper, max_iterations,
                       NUMERICS_DEFAULT if numerics is None else numerics)
    opt = NpyOptions(int(chunk_fits), int(in_flight), int(bool(uncertainties)), int(budget_bytes))
    gathered = None
    if consumer is None:
        out = [np.empty((n, p), np.float32), np.empty(n, np.int32), np.empty(n, np.uint32),
               np.empty(n, np.float32)]
        if uncertainties:
            out += [np.empty((n, p, p), np.float32), np.empty((n, p), np.float32),
                    np.empty(n, np.int32)]
        gathered = out
    failure = []

    def take(_user, first, count, params, status, iters, chi2, cov, se, us):
        try:
            def view(ptr, dtype, shape):
                size = int(np.prod(shape)) * np.dtype(dtype).itemsize
                buf = (ctypes.c_char * size).from_address(ptr)
                return np.frombuffer(buf, dtype).reshape(shape).copy()
            chunk = [view(params, np.float32, (count, p)), view(status, np.int32, (count,)),
                     view(iters, np.uint32, (count,)), view(chi2, np.float32, (count,))]
            if uncertainties:
                chunk += [view(cov, np.float32, (count, p, p)), view(se, np.float32, (count, p)),
                          view(us, np.int32, (count,))]
            if gathered is not None:
                for dst, src in zip(gathered, chunk):
                    dst[first:first + count] = src
            else:
                cls = FitResultWithUncertainties if uncertainties else FitResult
                consumer(int(first), cls(*chunk))
            return 0
        except BaseException as exc:  # reported after the pipeline has drained
            failure.append(exc)
            return 1

    callback = _CONSUMER(take)
    stats = NpyStats()
    err = _native.error_buffer()
    rc = lib.bf_hip_fit_npy(_native.MODELS[model], os.fsencode(x_path), os.fsencode(y_path),
                            None if p0_path is None else os.fsencode(p0_path), ctypes.byref(cfg),
                            ctypes.byref(opt), callback, None, ctypes.byref(stats), err, len(err))
    if failure:
        raise failure[0]
    _native.check(rc, err)
    if gathered is None:
        return None, stats.as_dict()
    cls = FitResultWithUncertainties if uncertainties else FitResult
    return cls(*gathered), stats.as_dict()
