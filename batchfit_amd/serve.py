"""Continuous-batching fit server (``bf_serve_*`` in include/batchfit_hip.h).

Requests are ``fit``-style batches (x[M], y[N,M], model, p0, bounds) submitted
from any thread; the native scheduler admits them in arrival order into free
device slots while the bytes the admitted batches hold stay within the HBM
budget, and runs each admitted group of requests as one ragged launch, so a
request that arrives while others run is picked up as soon as a slot and
memory free up. A request's result is the same as the same batch fitted alone
(every fit's arithmetic depends only on its own data, model and M).

The reference serialises native fitting behind one global mutex
(python/bindings.cpp:13,66) and pipelines only NPY chunks
(src/metal/pipeline.cpp:99-169); this server has no counterpart there.
"""
from __future__ import annotations

import ctypes
from typing import NamedTuple, Optional

import numpy as np

from . import _native


class ServeOptions(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("slots", ctypes.c_uint32),
                ("max_batch_fits", ctypes.c_uint32), ("numerics", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("rc", ctypes.c_int32), ("batch", ctypes.c_uint32), ("queued_s", ctypes.c_double),
                ("service_s", ctypes.c_double), ("latency_s", ctypes.c_double),
                ("message", ctypes.c_char * 256)]


class ServeCounters(ctypes.Structure):
    _fields_ = [("requests_done", ctypes.c_uint64), ("fits_done", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("pending", ctypes.c_uint64),
                ("bytes_in_flight", ctypes.c_uint64), ("max_bytes_in_flight", ctypes.c_uint64),
                ("max_batches_in_flight", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Timing(NamedTuple):
    """Seconds from submission to admission, admission to completion, and in total."""
    batch: int
    queued: float
    service: float
    latency: float


def _setup(lib):
    if getattr(lib, "_serve_ready", False):
        return
    vp, u32, i32, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t
    lib.bf_serve_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ServeOptions),
                                    ctypes.c_char_p, sz]
    lib.bf_serve_create.restype = i32
    lib.bf_serve_submit.argtypes = [vp, u32, ctypes.POINTER(_native.Config), i32] + [vp] * 10 + [
        ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, sz]
    lib.bf_serve_submit.restype = i32
    lib.bf_serve_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(RequestInfo)]
    lib.bf_serve_wait.restype = i32
    lib.bf_serve_stats.argtypes = [vp, ctypes.POINTER(ServeCounters)]
    lib.bf_serve_stats.restype = i32
    lib.bf_serve_destroy.argtypes = [vp]
    lib.bf_serve_destroy.restype = i32
    lib._serve_ready = True


class Ticket:
    """A submitted request: holds its input and output arrays until ``wait``."""

    def __init__(self, server, number, model, keep, result):
        self.server, self.number, self.model = server, number, model
        self._keep, self._result = keep, result
        self.timing: Optional[Timing] = None
        self.done = False

    def wait(self, timeout: Optional[float] = None):
        return self.server.wait(self, timeout)


class Server:
    """``Server(hbm_budget=…, slots=3, max_batch_fits=…, numerics=5)``.

    ``submit(x, y, model, p0, bounds, uncertainties=False)`` takes host arrays
    and returns a Ticket; ``submit_device(…)`` takes ``DeviceArray`` inputs and
    outputs (no uncertainties). ``wait(ticket)`` returns the FitResult (or
    FitResultWithUncertainties) and stores the request's latency split in
    ``ticket.timing``.
    """

    def __init__(self, hbm_budget: int = 0, slots: int = 0, max_batch_fits: int = 0,
                 numerics: Optional[int] = None):
        from . import NUMERICS_DEFAULT
        self.lib = _native.library()
        _setup(self.lib)
        opt = ServeOptions(int(hbm_budget), int(slots), int(max_batch_fits),
                           NUMERICS_DEFAULT if numerics is None else int(numerics), 0)
        self.handle = ctypes.c_void_p()
        err = _native.error_buffer()
        _native.check(self.lib.bf_serve_create(ctypes.byref(self.handle), ctypes.byref(opt), err,
                                               len(err)), err)

    # -- submission -------------------------------------------------------
    def _submit(self, model, cfg, on_device, ptrs, keep, result) -> Ticket:
        number = ctypes.c_uint64()
        err = _native.error_buffer()
        _native.check(self.lib.bf_serve_submit(self.handle, _native.MODELS[model],
                                               ctypes.byref(cfg), int(on_device), *ptrs,
                                               ctypes.byref(number), err, len(err)), err)
        return Ticket(self, number.value, model, keep, result)

    def submit(self, x, y, model="lorentzian", p0=None, bounds=None, *, uncertainties=False,
               max_iterations=100) -> Ticket:
        from . import FitResult, FitResultWithUncertainties, _make_config, _validate_batch
        x, y, p0, lower, upper = _validate_batch(x, y, model, p0, bounds)
        n, p = y.shape[0], len(lower)
        cfg = _make_config(n, x.size, lower, upper, max_iterations, 0)
        out = [np.empty((n, p), np.float32), np.empty(n, np.int32), np.empty(n, np.uint32),
               np.empty(n, np.float32)]
        if uncertainties:
            out += [np.empty((n, p, p), np.float32), np.empty((n, p), np.float32),
                    np.empty(n, np.int32)]
            result = FitResultWithUncertainties(*out)
        else:
            result = FitResult(*out)

        def ptr(a):
            return None if a is None else a.ctypes.data

        ptrs = [ptr(x), ptr(y), ptr(p0)] + [ptr(a) for a in out] + [None] * (7 - len(out))
        return self._submit(model, cfg, False, ptrs, (x, y, p0, cfg), result)

    def submit_device(self, x, y, model, n: int, m: int, out: dict, p0=None, bounds=None, *,
                      max_iterations=100) -> Ticket:
        """Device-resident request: ``x``/``y``/``p0`` are DeviceArrays (y[N,M],
        p0[N,P]); ``out`` maps params[N,P]/status/iterations/chi2 to DeviceArrays."""
        from . import _PARAMETERS, _make_config, _validate_batch
        p = _PARAMETERS[model]
        _, _, _, lower, upper = _validate_batch(np.zeros(m, np.float32),
                                                np.zeros((0, m), np.float32), model, None, bounds)
        cfg = _make_config(n, m, lower, upper, max_iterations, 0)
        need = {"params": n * p * 4, "status": n * 4, "iterations": n * 4, "chi2": n * 4}
        for key, size in need.items():
            if out[key].nbytes < size:
                raise ValueError(f"output {key} needs {size} bytes")
        if x.nbytes < m * 4 or y.nbytes < n * m * 4 or (p0 is not None and p0.nbytes < n * p * 4):
            raise ValueError("device inputs are smaller than x[M], y[N,M] and p0[N,P]")
        ptrs = [x.ptr, y.ptr, None if p0 is None else p0.ptr, out["params"].ptr,
                out["status"].ptr, out["iterations"].ptr, out["chi2"].ptr, None, None, None]
        return self._submit(model, cfg, True, ptrs, (x, y, p0, out, cfg), out)

    # -- completion -------------------------------------------------------
    def wait(self, ticket: Ticket, timeout: Optional[float] = None):
        """Block until the request is done; raise its error, if any."""
        if ticket.done:
            return ticket._result
        info = RequestInfo()
        rc = self.lib.bf_serve_wait(self.handle, ticket.number,
                                    -1.0 if timeout is None else float(timeout),
                                    ctypes.byref(info))
        if rc == _native.BF_TIMEOUT:
            raise TimeoutError(f"request {ticket.number} not finished within {timeout} s")
        ticket.done = True
        ticket.timing = Timing(info.batch, info.queued_s, info.service_s, info.latency_s)
        ticket._keep = None
        if rc != _native.BF_OK:
            buffer = ctypes.create_string_buffer(info.message)
            _native.check(rc, buffer)
        return ticket._result

    def stats(self) -> dict:
        c = ServeCounters()
        self.lib.bf_serve_stats(self.handle, ctypes.byref(c))
        return {name: getattr(c, name) for name, _ in ServeCounters._fields_ if name != "reserved"}

    def close(self) -> None:
        """Finish every submitted request, then stop the server."""
        if self.handle.value:
            self.lib.bf_serve_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
