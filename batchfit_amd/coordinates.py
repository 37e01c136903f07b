"""Coordinate normalisation for fits whose units are far from 1
(include/batchfit/coordinates.hpp of the reference, NUMERICS.md:9-20).

FP32 fitting loses the sample positions of a grid such as x = 1e6 ± 0.005 and
the scale of signals such as 1e-6 before any arithmetic starts. The remedy is
host arithmetic in float64: map x to [-1, 1] around the grid's midpoint and
divide y by its largest magnitude, fit the normalised float32 arrays on the
device, and map parameters, χ² and covariance back in float64. Bounds and p0
passed to the fit are in the fitted (normalised) units.

``Coordinates`` holds one frame (origin, x_scale, y_scale) and the maps of the
reference for the four models: parameters (``normalize`` / ``physical``),
χ² (``physical_chi_squared``) and covariance (``physical_covariance``, the
congruence T·C·Tᵀ with T the Jacobian of the physical parameters with
respect to the normalised ones). Every map also takes stacked rows ([..., P]
or [..., P, P]) with y_scale broadcast against the leading axes, which is how
``fit_normalized`` gives each spectrum of a batch its own signal scale over
the shared grid.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Union

import numpy as np

_P = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8}


def _count(model: str) -> int:
    if model not in _P:
        raise ValueError(f"unknown model {model!r}; choose one of {tuple(_P)}")
    return _P[model]


def _midpoint(a: float, b: float) -> float:
    """std::midpoint of two finite doubles (the libstdc++ branches: (a+b)/2
    unless that could overflow or lose a subnormal half)."""
    big = np.finfo(np.float64).max / 2
    small = np.finfo(np.float64).tiny * 2
    if abs(a) <= big and abs(b) <= big:
        return (a + b) / 2
    if abs(a) < small:
        return a + b / 2
    if abs(b) < small:
        return a / 2 + b
    return a / 2 + b / 2


@dataclass(frozen=True)
class Coordinates:
    """A frame x → (x − origin)/x_scale, y → y/y_scale (float64).
    ``y_scale`` may be an array (one scale per spectrum of a batch)."""
    origin: float
    x_scale: float
    y_scale: Union[float, np.ndarray]

    @staticmethod
    def _grid(x):
        x = np.asarray(x, np.float64)
        if x.ndim != 1 or not x.size:
            raise ValueError("normalization requires matching nonempty x/y")
        if not np.all(np.isfinite(x)):
            raise ValueError("normalization requires finite inputs")
        lo, hi = float(x.min()), float(x.max())
        half = hi / 2 - lo / 2  # no overflow for grids near ±max
        if not (half > 0 and np.isfinite(half)):
            raise ValueError("normalization requires a resolved x range")
        return _midpoint(lo, hi), half

    @classmethod
    def from_data(cls, x, y) -> "Coordinates":
        """The frame of one spectrum (Coordinates::from, coordinates.hpp:17-31):
        origin = midpoint of the x range, x_scale = half of it, y_scale = the
        largest |y| (1 for an all-zero y)."""
        y = np.asarray(y, np.float64)
        if y.ndim != 1 or y.size != np.asarray(x).size:
            raise ValueError("normalization requires matching nonempty x/y")
        if not np.all(np.isfinite(y)):
            raise ValueError("normalization requires finite inputs")
        origin, half = cls._grid(x)
        signal = float(np.max(np.abs(y)))
        return cls(origin, half, signal if signal > 0 else 1.0)

    @classmethod
    def from_batch(cls, x, y) -> "Coordinates":
        """One frame per spectrum of y[N, M] over the shared x[M]: the same
        origin and x_scale for all, y_scale[N] from each row (each row's frame
        equals from_data(x, y[i]))."""
        y = np.asarray(y, np.float64)
        if y.ndim != 2 or y.shape[1] != np.asarray(x).size:
            raise ValueError("normalization requires matching nonempty x/y")
        if not np.all(np.isfinite(y)):
            raise ValueError("normalization requires finite inputs")
        origin, half = cls._grid(x)
        signal = np.max(np.abs(y), axis=1) if y.shape[0] else np.zeros(0)
        return cls(origin, half, np.where(signal > 0, signal, 1.0))

    def _ys(self, extra_dims: int):
        ys = np.asarray(self.y_scale, np.float64)
        return ys.reshape(ys.shape + (1,) * extra_dims)

    def x(self, value):
        return (np.asarray(value, np.float64) - self.origin) / self.x_scale

    def y(self, value):
        value = np.asarray(value, np.float64)
        return value / self._ys(value.ndim - np.ndim(self.y_scale))

    def physical_chi_squared(self, value):
        """χ² of the normalised fit in the physical y units (χ²·y_scale²)."""
        ys = np.asarray(self.y_scale, np.float64)
        return (np.asarray(value, np.float64) * ys) * ys

    def normalize(self, model: str, params) -> np.ndarray:
        """Physical parameters [..., P] → the parameters of the same curve in
        the normalised frame (the affine baseline b + s·x of the two-peak
        model becomes b' + s'·u with u the normalised x)."""
        p = _count(model)
        q = np.array(params, np.float64, copy=True)
        if q.shape[-1:] != (p,):
            raise ValueError("normalization parameter count mismatch")
        src = np.asarray(params, np.float64)
        ys = np.asarray(self.y_scale, np.float64)
        q[..., 0] = src[..., 0] / ys
        q[..., 1] = (src[..., 1] - self.origin) / self.x_scale
        q[..., 2] = src[..., 2] / self.x_scale
        if p == 8:
            q[..., 3] = src[..., 3] / ys
            q[..., 4] = (src[..., 4] - self.origin) / self.x_scale
            q[..., 5] = src[..., 5] / self.x_scale
            q[..., 6] = (src[..., 7] * self.origin + src[..., 6]) / ys
            q[..., 7] = src[..., 7] * self.x_scale / ys
        return q

    def physical(self, model: str, params) -> np.ndarray:
        """Normalised parameters [..., P] → physical ones (the inverse of normalize)."""
        p = _count(model)
        src = np.asarray(params, np.float64)
        if src.shape[-1:] != (p,):
            raise ValueError("physical parameter count mismatch")
        ys = np.asarray(self.y_scale, np.float64)
        r = np.array(src, copy=True)
        r[..., 0] = src[..., 0] * ys
        r[..., 1] = src[..., 1] * self.x_scale + self.origin
        r[..., 2] = src[..., 2] * self.x_scale
        if p == 8:
            r[..., 3] = src[..., 3] * ys
            r[..., 4] = src[..., 4] * self.x_scale + self.origin
            r[..., 5] = src[..., 5] * self.x_scale
            r[..., 7] = src[..., 7] * ys / self.x_scale
            r[..., 6] = src[..., 6] * ys - r[..., 7] * self.origin
        return r

    def transform(self, model: str) -> np.ndarray:
        """T[..., P, P] = ∂ physical / ∂ normalised (linear maps, so exact)."""
        p = _count(model)
        ys = np.asarray(self.y_scale, np.float64)
        t = np.zeros(ys.shape + (p, p))
        t[..., 0, 0] = ys
        t[..., 1, 1] = t[..., 2, 2] = self.x_scale
        if p == 4:
            t[..., 3, 3] = 1.0
        if p == 8:
            t[..., 3, 3] = ys
            t[..., 4, 4] = t[..., 5, 5] = self.x_scale
            t[..., 6, 6] = ys
            t[..., 6, 7] = -self.origin * ys / self.x_scale
            t[..., 7, 7] = ys / self.x_scale
        return t

    def physical_covariance(self, model: str, covariance) -> np.ndarray:
        """Covariance of the normalised parameters [..., P, P] → physical units, T·C·Tᵀ."""
        p = _count(model)
        c = np.asarray(covariance, np.float64)
        if c.shape[-2:] != (p, p):
            raise ValueError("covariance shape mismatch")
        t = self.transform(model)
        return t @ c @ np.swapaxes(t, -1, -2)


def fit_normalized(x, y, model="lorentzian", p0=None, bounds=None, *, uncertainties=False,
                   **options):
    """The reference's recipe for difficult units (NUMERICS.md:9-20): normalise
    x[M] and y[N, M] in float64 (one y scale per spectrum), fit the float32
    normalised arrays on the device with compensated partials and the scaled
    Cholesky solve, and map parameters, χ² and covariance back in float64.
    ``p0`` is in physical units (mapped with each row's frame); ``bounds`` are
    in the fitted, normalised units, as in the reference. Returns the
    FitResult fields in float64 physical units, plus the frame."""
    from . import FitResult, FitResultWithUncertainties, fit
    p = _count(model)
    x64 = np.asarray(x, np.float64)
    y64 = np.asarray(y, np.float64)
    if y64.ndim == 1:
        y64 = y64.reshape(1, -1)
    frame = Coordinates.from_batch(x64, y64)
    u = frame.x(x64).astype(np.float32)
    v = frame.y(y64).astype(np.float32)
    q0 = None
    if p0 is not None:
        q0 = frame.normalize(model, np.broadcast_to(np.asarray(p0, np.float64),
                                                    (y64.shape[0], p))).astype(np.float32)
    options.setdefault("numerics", 7)  # compensated + scaled Cholesky + FP64 solve
    r = fit(u, v, model, q0, bounds, "hip", uncertainties=uncertainties, **options)
    params = frame.physical(model, r.params.astype(np.float64))
    chi2 = frame.physical_chi_squared(r.chi2.astype(np.float64))
    if not uncertainties:
        return FitResult(params, r.status, r.iterations, chi2), frame
    cov = frame.physical_covariance(model, r.covariance.astype(np.float64))
    se = np.sqrt(np.maximum(np.diagonal(cov, axis1=-2, axis2=-1), 0.0))
    se = np.where(np.isnan(r.stderr), np.nan, se)
    return FitResultWithUncertainties(params, r.status, r.iterations, chi2, cov, se,
                                      r.uncertainty_status), frame
