"""Ragged batches: fits of mixed model, M and bounds in one captured launch.

``fit_ragged(batches)`` packs several uniform batches — each ``(x, y, model,
p0, bounds)`` as accepted by ``batchfit_amd.fit`` — into one plan
(``bf_hip_ragged_create``) and runs it once on the device. Each fit is solved
with the mapping its own (model, M) selects, so results do not depend on the
other fits in the plan.

``RaggedPlan`` is the thin wrapper over a plan for callers that keep their
data in device memory (``DeviceArray``), as the serving loop and bench.py do.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import _native

FIT_DESC = np.dtype([("y_offset", "<u8"), ("points", "<u4"), ("grid", "<u4"), ("model", "<u4"),
                     ("config", "<u4"), ("p0_row", "<u4"), ("out_row", "<u4")])
GRID_DESC = np.dtype([("x_offset", "<u8"), ("points", "<u4"), ("reserved", "<u4")])
NO_P0 = 0xFFFFFFFF
assert FIT_DESC.itemsize == 32 and GRID_DESC.itemsize == 16


class DeviceArray:
    """A device buffer owned by Python (hipMalloc through the library)."""

    def __init__(self, nbytes: int, host: Optional[np.ndarray] = None):
        self.lib = _native.library()
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        self._owner = True
        self._base = None
        err = _native.error_buffer()
        _native.check(self.lib.bf_hip_malloc(ctypes.byref(self.ptr), max(1, self.nbytes), err,
                                             len(err)), err)
        if host is not None:
            self.upload(host)

    def view(self, offset: int, nbytes: int) -> "DeviceArray":
        """A non-owning window [offset, offset+nbytes) of this buffer."""
        if offset < 0 or nbytes < 0 or offset + nbytes > self.nbytes:
            raise ValueError("view exceeds the device buffer")
        v = object.__new__(DeviceArray)
        v.lib, v.nbytes, v._owner, v._base = self.lib, int(nbytes), False, self
        v.ptr = ctypes.c_void_p(self.ptr.value + int(offset))
        return v

    @classmethod
    def of(cls, host: np.ndarray) -> "DeviceArray":
        host = np.ascontiguousarray(host)
        return cls(host.nbytes, host)

    def upload(self, host: np.ndarray, offset: int = 0) -> None:
        host = np.ascontiguousarray(host)
        if offset + host.nbytes > self.nbytes:
            raise ValueError("upload exceeds the device buffer")
        err = _native.error_buffer()
        dst = ctypes.c_void_p(self.ptr.value + offset)
        _native.check(self.lib.bf_hip_memcpy(dst, host.ctypes.data, host.nbytes, 1, err,
                                             len(err)), err)

    def download(self, host: np.ndarray, offset: int = 0) -> np.ndarray:
        if offset + host.nbytes > self.nbytes:
            raise ValueError("download exceeds the device buffer")
        err = _native.error_buffer()
        src = ctypes.c_void_p(self.ptr.value + offset)
        _native.check(self.lib.bf_hip_memcpy(host.ctypes.data, src, host.nbytes, 2, err,
                                             len(err)), err)
        return host

    def free(self) -> None:
        if self._owner and self.ptr.value:
            self.lib.bf_hip_free(self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _setup(lib):
    if getattr(lib, "_ragged_ready", False):
        return
    vp, u32, i32, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t
    lib.bf_hip_ragged_create.argtypes = [ctypes.POINTER(vp), u32, vp, u32, vp, u32, vp, u32, i32,
                                         vp, ctypes.c_char_p, sz]
    lib.bf_hip_ragged_create.restype = i32
    lib.bf_hip_ragged_run.argtypes = [vp] * 11 + [ctypes.c_char_p, sz]
    lib.bf_hip_ragged_run.restype = i32
    lib.bf_hip_ragged_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double), i32]
    lib.bf_hip_ragged_timing.restype = i32
    lib.bf_hip_ragged_destroy.argtypes = [vp]
    lib.bf_hip_ragged_destroy.restype = i32
    lib._ragged_ready = True


class RaggedPlan:
    """An uploaded ragged batch (descriptor tables, class order, work queues)."""

    def __init__(self, fits: np.ndarray, grids: np.ndarray, configs: Sequence[_native.Config],
                 numerics: Optional[int] = None, uncertainties: bool = False, stream=None):
        self.lib = _native.library()
        _setup(self.lib)
        fits = np.ascontiguousarray(fits, FIT_DESC)
        grids = np.ascontiguousarray(grids, GRID_DESC)
        table = (_native.Config * max(1, len(configs)))(*configs)
        self.fits = len(fits)
        self.rows = int(fits["out_row"].max()) + 1 if len(fits) else 0
        self.uncertainties = bool(uncertainties)
        self.handle = ctypes.c_void_p()
        err = _native.error_buffer()
        numerics = 5 if numerics is None else int(numerics)
        _native.check(self.lib.bf_hip_ragged_create(
            ctypes.byref(self.handle), len(fits), fits.ctypes.data, len(grids), grids.ctypes.data,
            len(configs), ctypes.cast(table, ctypes.c_void_p), numerics, int(uncertainties),
            stream, err, len(err)), err)

    def run(self, x: DeviceArray, y: DeviceArray, p0: Optional[DeviceArray], out: dict) -> None:
        """Enqueue on the plan's stream. ``out`` maps params/status/iterations/chi2
        (and cov/stderr/ustatus) to DeviceArrays sized for ``self.rows``."""
        err = _native.error_buffer()

        def ptr(a):
            return None if a is None else a.ptr

        _native.check(self.lib.bf_hip_ragged_run(
            self.handle, x.ptr, y.ptr, ptr(p0), out["params"].ptr, out["status"].ptr,
            out["iterations"].ptr, out["chi2"].ptr, ptr(out.get("cov")), ptr(out.get("stderr")),
            ptr(out.get("ustatus")), err, len(err)), err)

    def timing(self):
        times = (ctypes.c_double * 2)()
        self.lib.bf_hip_ragged_timing(self.handle, times, 2)
        return list(times)

    def close(self) -> None:
        if self.handle.value:
            self.lib.bf_hip_ragged_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def allocate_outputs(rows: int, uncertainties: bool) -> dict:
    out = {"params": DeviceArray(rows * 32), "status": DeviceArray(rows * 4),
           "iterations": DeviceArray(rows * 4), "chi2": DeviceArray(rows * 4)}
    if uncertainties:
        out.update(cov=DeviceArray(rows * 256), stderr=DeviceArray(rows * 32),
                   ustatus=DeviceArray(rows * 4))
    return out


def download_outputs(out: dict, rows: int) -> dict:
    host = {"params": np.empty((rows, 8), np.float32), "status": np.empty(rows, np.int32),
            "iterations": np.empty(rows, np.uint32), "chi2": np.empty(rows, np.float32)}
    if "cov" in out:
        host.update(cov=np.empty((rows, 8, 8), np.float32), stderr=np.empty((rows, 8), np.float32),
                    ustatus=np.empty(rows, np.int32))
    for key, array in host.items():
        out[key].download(array)
    return host


def fit_ragged(batches: Sequence, *, uncertainties: bool = False, numerics: Optional[int] = None,
               max_iterations: int = 100) -> List:
    """Fit several batches ``(x[M], y[N,M] or y[M], model, p0, bounds)`` in one
    ragged launch; returns one FitResult (or FitResultWithUncertainties) per batch."""
    from . import (FitResult, FitResultWithUncertainties, NUMERICS_DEFAULT, _make_config,
                   _validate_batch)
    if not batches:
        return []
    numerics = NUMERICS_DEFAULT if numerics is None else int(numerics)
    xs, ys, p0s, descs, grids, configs, shapes = [], [], [], [], [], [], []
    x_off = y_off = row = p0_rows = 0
    for b, (x, y, model, p0, bounds) in enumerate(batches):
        x, y, p0, lower, upper = _validate_batch(x, y, model, p0, bounds)
        n, m, p = y.shape[0], x.size, len(lower)
        cfg = _make_config(n, m, lower, upper, max_iterations, numerics)
        configs.append(cfg)
        grids.append((x_off, m, 0))
        xs.append(x)
        d = np.zeros(n, FIT_DESC)
        d["y_offset"] = y_off + np.arange(n, dtype=np.uint64) * m
        d["points"], d["grid"], d["model"], d["config"] = m, b, _native.MODELS[model], b
        d["out_row"] = row + np.arange(n, dtype=np.uint32)
        if p0 is None:
            d["p0_row"] = NO_P0
        else:
            q = np.zeros((n, 8), np.float32)
            q[:, :p] = np.broadcast_to(np.asarray(p0, np.float32), (n, p))
            p0s.append(q)
            d["p0_row"] = p0_rows + np.arange(n, dtype=np.uint32)
            p0_rows += n
        descs.append(d)
        ys.append(np.ascontiguousarray(y).reshape(-1))
        shapes.append((row, n, p))
        x_off += m
        y_off += n * m
        row += n
    fits = np.concatenate(descs)
    plan = RaggedPlan(fits, np.array(grids, GRID_DESC), configs, numerics, uncertainties)
    try:
        dx = DeviceArray.of(np.concatenate(xs))
        dy = DeviceArray.of(np.concatenate(ys))
        dp0 = DeviceArray.of(np.concatenate(p0s)) if p0s else None
        out = allocate_outputs(max(1, row), uncertainties)
        if row:
            plan.run(dx, dy, dp0, out)
        host = download_outputs(out, max(1, row))
    finally:
        plan.close()
    results = []
    for r0, n, p in shapes:
        s = slice(r0, r0 + n)
        base = (host["params"][s, :p].copy(), host["status"][s].copy(),
                host["iterations"][s].copy(), host["chi2"][s].copy())
        if uncertainties:
            results.append(FitResultWithUncertainties(
                *base, host["cov"].reshape(-1, 64)[s, :p * p].reshape(n, p, p).copy(),
                host["stderr"][s, :p].copy(),
                host["ustatus"][s].copy()))
        else:
            results.append(FitResult(*base))
    return results
