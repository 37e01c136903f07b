"""Seeded synthetic spectra generated directly in HBM (``bf_hip_synthesize``).

Each fit i draws its parameters uniformly from [lower, upper] and its noise
from Philox4x32-10 counters keyed by (seed, fit_offset + i), so a batch can be
generated in pieces and the values do not depend on the launch.
``layout="doublet"`` (two_lorentzian) draws one HWHM for both lines and places
the second centre at c1 + U[sep]·HWHM. Noise modes: ``"absolute"`` (σ =
noise), ``"relative"`` (σ = noise·amplitude), and the reference generator's
``"snr_gaussian"`` (σ = RMS of the clean row / snr) and ``"snr_poisson"``
(counts of mean gain·f, gain = snr²·Σf/Σf², divided by the gain)
(src/cpu/synthetic.cpp:83-108); for the SNR modes ``noise`` is the snr.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _native
from .ragged import DeviceArray

_LAYOUTS = {"independent": 0, "doublet": 1}
NOISE_MODES = {"absolute": 0, "relative": 1, "snr_gaussian": 2, "snr_poisson": 3}


def _setup(lib):
    if getattr(lib, "_synth_ready", False):
        return
    vp, u32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float
    lib.bf_hip_synthesize.argtypes = [u32, u32, u32, vp, vp, vp, u32, f32, f32, u32, f32,
                                      ctypes.c_uint64, ctypes.c_uint64, vp, vp, vp,
                                      ctypes.c_char_p, ctypes.c_size_t]
    lib.bf_hip_synthesize.restype = ctypes.c_int
    lib._synth_ready = True


def synthesize(model: str, n: int, x: DeviceArray, m: int, lower: Sequence[float],
               upper: Sequence[float], y: DeviceArray, *, noise: float = 0.0,
               noise_relative: bool = False, seed: int = 0, fit_offset: int = 0,
               layout: str = "independent", separation=(2.0, 6.0),
               truth: Optional[DeviceArray] = None, stream=None,
               noise_mode: Optional[str] = None) -> None:
    """Write y[n, m] (and truth[n, 8] when given) on the library stream.
    ``noise_mode`` (see the module docstring) overrides ``noise_relative``."""
    lib = _native.library()
    _setup(lib)
    if model not in _native.MODELS or layout not in _LAYOUTS:
        raise ValueError("unknown model or layout")
    if noise_mode is None:
        noise_mode = "relative" if noise_relative else "absolute"
    if noise_mode not in NOISE_MODES:
        raise ValueError(f"noise_mode must be one of {tuple(NOISE_MODES)}")
    if y.nbytes < n * m * 4 or x.nbytes < m * 4 or (truth is not None and truth.nbytes < n * 32):
        raise ValueError("device buffers are smaller than x[m], y[n,m] and truth[n,8]")
    lo = np.zeros(8, np.float32)
    hi = np.zeros(8, np.float32)
    p = _native.PARAMETERS[model]
    lo[:p], hi[:p] = np.asarray(lower, np.float32)[:p], np.asarray(upper, np.float32)[:p]
    err = _native.error_buffer()
    _native.check(lib.bf_hip_synthesize(
        _native.MODELS[model], int(n), int(m), x.ptr, lo.ctypes.data, hi.ctypes.data,
        _LAYOUTS[layout], float(separation[0]), float(separation[1]), NOISE_MODES[noise_mode],
        float(noise), int(seed), int(fit_offset), y.ptr, None if truth is None else truth.ptr,
        stream, err, len(err)), err)
