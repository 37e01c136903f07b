"""Seeded synthetic inputs shared by the golden-vector generator, the tests,
``__graft_entry__.smoke`` and ``bench.py``.

Generators use NumPy's PCG64 ``default_rng``, whose streams are stable across
NumPy versions. Noise rows are drawn after all parameter arrays, in row
order, so a chunked draw equals one draw of the whole array.
"""
from __future__ import annotations

import numpy as np

LN2 = np.log(2.0)


def lorentzian(x, a, c, w):
    return a / (1 + ((x - c) / w) ** 2)


def gaussian(x, a, c, s):
    return a * np.exp(-0.5 * ((x - c) / s) ** 2)


def pseudo_voigt(x, a, c, f, eta):
    z = (x - c) / f
    return a * ((1 - eta) * np.exp(-4 * LN2 * z * z) + eta / (1 + 4 * z * z))


def two_lorentzian(x, a1, c1, w1, a2, c2, w2, b0, b1):
    return lorentzian(x, a1, c1, w1) + lorentzian(x, a2, c2, w2) + b0 + b1 * x


MODELS = {"gaussian": gaussian, "lorentzian": lorentzian, "pseudo_voigt": pseudo_voigt,
          "two_lorentzian": two_lorentzian}


def config1(dtype=np.float64):
    """README quickstart (README.md:79-91 of the reference): Lorentzian, N=1000,
    M=129, x=linspace(-5,5), a=2, HWHM 0.7, centres linspace(-0.5,0.5),
    noise 0.02·N(0,1) from default_rng(123); bounds [0,-5,1e-4]–[inf,5,5]."""
    dtype = np.dtype(dtype)
    x = np.linspace(-5, 5, 129, dtype=dtype)
    centers = np.linspace(-0.5, 0.5, 1000, dtype=dtype)
    y = 2 / (1 + ((x - centers[:, None]) / 0.7) ** 2)
    rng = np.random.default_rng(123)
    y = y + 0.02 * rng.standard_normal(y.shape, dtype)
    bounds = (np.array([0, -5, 1e-4]), np.array([np.inf, 5, 5]))
    return x, y.astype(dtype), bounds


def pv_params(n, seed=7):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 5, n)
    c = rng.uniform(-3, 3, n)
    f = rng.uniform(0.2, 2, n)
    eta = rng.uniform(0, 1, n)
    return rng, np.column_stack([a, c, f, eta])


def config2(n=1 << 20, m=1024, seed=7, rows=None, chunk=1 << 16):
    """Metric workload: pseudo-Voigt P=4, x=linspace(-10,10,M) float32;
    a~U[0.5,5], c~U[-3,3], FWHM~U[0.2,2], η~U[0,1]; Gaussian noise σ=1% of
    amplitude; bounds [0,-10,1e-4,0]–[inf,10,10,1]. ``rows`` limits the
    returned rows to the first ``rows`` of the same stream."""
    rng, truth = pv_params(n, seed)
    x64 = np.linspace(-10, 10, m)
    x = x64.astype(np.float32)
    rows = n if rows is None else min(rows, n)
    y = np.empty((rows, m), np.float32)
    for lo in range(0, rows, chunk):
        hi = min(rows, lo + chunk)
        t = truth[lo:hi]
        z = (x64[None, :] - t[:, 1:2]) / t[:, 2:3]
        clean = t[:, 0:1] * ((1 - t[:, 3:4]) * np.exp(-4 * LN2 * z * z) + t[:, 3:4] / (1 + 4 * z * z))
        noise = rng.standard_normal((hi - lo, m))
        y[lo:hi] = clean + noise * (0.01 * t[:, 0:1])
    bounds = (np.array([0, -10, 1e-4, 0]), np.array([np.inf, 10, 10, 1]))
    return x, y, bounds, truth[:rows]


def config3(n=262144, m=2048, seed=11, rows=None):
    """Overlapping doublets: two_lorentzian (P=8 with intercept and slope fixed
    to 0 by equal bounds), x=linspace(-20,20,M) float32; a1,a2~U[0.5,5],
    c1~U[-5,5], one HWHM~U[0.2,1.5] for both peaks, c2=c1+U[0.5,3]·HWHM,
    noise σ=0.01; 3% rows constant, 2% rows with one NaN."""
    rng = np.random.default_rng(seed)
    a1 = rng.uniform(0.5, 5, n)
    a2 = rng.uniform(0.5, 5, n)
    c1 = rng.uniform(-5, 5, n)
    w = rng.uniform(0.2, 1.5, n)
    c2 = c1 + rng.uniform(0.5, 3, n) * w
    kind = rng.uniform(0, 1, n)
    nan_at = rng.integers(0, m, n)
    x64 = np.linspace(-20, 20, m)
    rows = n if rows is None else min(rows, n)
    y = np.empty((rows, m), np.float32)
    for i0 in range(0, rows, 8192):
        i1 = min(rows, i0 + 8192)
        s = slice(i0, i1)
        clean = (lorentzian(x64[None], a1[s, None], c1[s, None], w[s, None])
                 + lorentzian(x64[None], a2[s, None], c2[s, None], w[s, None]))
        y[s] = clean + 0.01 * rng.standard_normal((i1 - i0, m))
    constant = kind[:rows] < 0.03
    with_nan = (kind[:rows] >= 0.03) & (kind[:rows] < 0.05)
    y[constant] = y[constant, :1]
    y[np.nonzero(with_nan)[0], nan_at[:rows][with_nan]] = np.nan
    lower = np.array([0, -20, 1e-4, 0, -20, 1e-4, 0, 0])
    upper = np.array([np.inf, 20, 20, np.inf, 20, 20, 0, 0])
    truth = np.column_stack([a1, c1, w, a2, c2, w, np.zeros(n), np.zeros(n)])[:rows]
    return x64.astype(np.float32), y, (lower, upper), truth, constant, with_nan


def comparison_set(model, n=257, m=129, seed=20261109):
    """Per-model sets in the style of python/tests/conftest.py:37-62 of the
    reference: x=linspace(-5,5,M), truth drawn per fit, noise at SNR 30, fixed
    p0 and bounds."""
    rng = np.random.default_rng(seed + list(MODELS).index(model) + 7 * m)
    x = np.linspace(-5, 5, m)
    if model == "two_lorentzian":
        truth = np.column_stack([rng.uniform(1.5, 3, n), rng.uniform(-2, -1.2, n),
                                 rng.uniform(.4, .7, n), rng.uniform(1.5, 3, n),
                                 rng.uniform(1.2, 2, n), rng.uniform(.4, .7, n),
                                 rng.uniform(.2, .5, n), rng.uniform(-.02, .02, n)])
        p0 = np.array([2, -1.5, .6, 2, 1.5, .6, .3, 0])
        bounds = (np.array([0, -5, .01, 0, -5, .01, -np.inf, -np.inf]),
                  np.array([np.inf, 5, 5, np.inf, 5, 5, np.inf, np.inf]))
    else:
        truth = np.column_stack([rng.uniform(1.5, 3, n), rng.uniform(-.5, .5, n),
                                 rng.uniform(.5, 1, n)])
        p0 = np.array([2, 0, .8])
        bounds = (np.array([0, -5, .01]), np.array([np.inf, 5, 5]))
        if model == "pseudo_voigt":
            truth[:, 2] *= 2
            truth = np.column_stack([truth, rng.uniform(.2, .8, n)])
            p0 = np.array([2, 0, 1.6, .5])
            bounds = (np.r_[bounds[0], 0], np.r_[bounds[1], 1])
    clean = np.stack([MODELS[model](x, *t) for t in truth])
    deviation = np.sqrt(np.mean(clean ** 2, axis=1)) / 30
    y = clean + rng.normal(size=clean.shape) * deviation[:, None]
    return x, y, p0, bounds, truth


def isolation_fixture():
    """GPU failure-isolation fixture (tests/metal_lm_checks.hpp:59-128 of the
    reference): Lorentzian, M=3, x≡0, y≡2, initial rows {1,0,1 | 2,0,1 | 1,0,0};
    expected statuses [2,0,3] with iterations [1,0,0]."""
    x = np.zeros(3, np.float32)
    y = np.full((3, 3), 2, np.float32)
    p0 = np.array([[1, 0, 1], [2, 0, 1], [1, 0, 0]], np.float32)
    return x, y, p0


def pathology_fixture(model, dtype=np.float32):
    """Nine-row pathology matrix (tests/numerics_checks.hpp:280-340 of the
    reference): x=-1+j/64 (M=129); rows zero, flat 3, NaN/+inf/-inf at point 17,
    edge peak (c=-1), sub-spacing width 1e-4, ordinary, 1e30 amplitude."""
    p = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8}[model]
    if p == 8:
        base = np.array([2, -.4, .12, 1.5, .4, .15, .2, .01])
    elif p == 4:
        base = np.array([2, 0, .2, .4])
    else:
        base = np.array([2, 0, .2])
    base = base.astype(dtype)
    m = 129
    x = (-1 + np.arange(m, dtype=dtype) / dtype(64)).astype(dtype)
    inits = np.tile(base, (9, 1)).astype(dtype)
    inits[5, 1] = -1
    inits[6, 2] = 1e-4
    if p == 8:
        inits[6, 5] = 1e-4
    inits[8, 0] *= dtype(1e30)
    if p == 8:
        inits[8, 3] *= dtype(1e30)
    return x, inits, base


def api_isolation_fixture(dtype=np.float64):
    """Row isolation of python/tests/test_api.py:69-76: statuses [0,3,3,4,0]."""
    x = np.linspace(-4, 4, 129, dtype=dtype)
    y = np.tile(2 / (1 + ((x - .2) / .7) ** 2), (5, 1)).astype(dtype)
    y[1, 10] = np.nan
    y[2, 10] = np.inf
    y[3] = 0
    p0 = np.tile(np.array([1.8, 0, .8], dtype=dtype), (5, 1))
    return x, y, p0
