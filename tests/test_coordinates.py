"""Coordinate normalisation (batchfit_amd.coordinates; the reference's
include/batchfit/coordinates.hpp), host float64: the reference's own checks
(tests/numerics_checks.hpp:63-94, tests/uncertainty_checks.hpp:76-86), the
covariance map against a finite-difference Jacobian of ``physical``, and
argument errors. The GPU test fits far-from-unit data through the frame."""
import numpy as np
import pytest

import datasets
from batchfit_amd.coordinates import Coordinates, fit_normalized

P = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8}
BASE = {3: [2, 0, .2], 4: [2, 0, .2, .4], 8: [2, -.4, .12, 1.5, .4, .15, .2, .01]}


def physical_params(p, offset, scale, signal):
    """tests/numerics_checks.hpp:66-77: unit parameters in a frame far from 1."""
    q = np.array(BASE[p], np.float64)
    q[0] *= signal
    q[1] = q[1] * scale + offset
    q[2] *= scale
    if p == 8:
        q[3] *= signal
        q[4] = q[4] * scale + offset
        q[5] *= scale
        q[7] *= signal / scale
        q[6] = -q[7] * offset + q[6] * signal
    return q


def model64(model, x, q):
    return datasets.MODELS[model](x, *q)


@pytest.mark.parametrize("model", list(P))
def test_round_trip_and_model_equivalence(model):
    p = P[model]
    offset, scale, signal = 1e6, .01, 1e-6
    phys = physical_params(p, offset, scale, signal)
    x = offset + scale * (-1 + np.arange(129) / 64)
    y = model64(model, x, phys)
    frame = Coordinates.from_data(x, y)
    q = frame.normalize(model, phys)
    back = frame.physical(model, q)
    np.testing.assert_allclose(back, phys, rtol=1e-12, atol=1e-14)
    u, v = frame.x(x), frame.y(y)
    np.testing.assert_allclose(model64(model, u, q), v, rtol=1e-7, atol=1e-10)
    assert abs(u[0] + 1) < 1e-7 and abs(u[-1] - 1) < 1e-7
    assert frame.y_scale == np.max(np.abs(y))


def test_physical_covariance_known_values():
    """tests/uncertainty_checks.hpp:76-86: frame {5, 2, 3}, identity with a
    0.4 intercept/slope correlation → 47.25, -9.45, 2.25."""
    frame = Coordinates(5.0, 2.0, 3.0)
    c = np.eye(8)
    c[6, 7] = c[7, 6] = .4
    out = frame.physical_covariance("two_lorentzian", c)
    assert out[6, 6] == pytest.approx(47.25, abs=1e-12)
    assert out[6, 7] == pytest.approx(-9.45, abs=1e-12)
    assert out[7, 7] == pytest.approx(2.25, abs=1e-12)


@pytest.mark.parametrize("model", list(P))
def test_covariance_map_is_the_jacobian_of_physical(model):
    """T·C·Tᵀ with T = ∂physical/∂normalised, T by central differences of
    physical() (a wrong sign or entry of the transform fails here)."""
    p = P[model]
    frame = Coordinates(3.5, 0.25, 7.0)
    q = np.array(BASE[p], np.float64)
    t = np.empty((p, p))
    for k in range(p):
        h = 1e-6
        d = np.zeros(p)
        d[k] = h
        t[:, k] = (frame.physical(model, q + d) - frame.physical(model, q - d)) / (2 * h)
    rng = np.random.default_rng(5)
    a = rng.normal(size=(p, p))
    c = a @ a.T
    np.testing.assert_allclose(frame.physical_covariance(model, c), t @ c @ t.T, rtol=1e-7,
                               atol=1e-9)


def test_batch_frames_equal_single_frames():
    x = np.linspace(-3, 7, 50)
    y = np.random.default_rng(2).normal(size=(6, 50)) * np.array([1, 1e-3, 1e5, 2, 0, 3])[:, None]
    batch = Coordinates.from_batch(x, y)
    for i in range(6):
        one = Coordinates.from_data(x, y[i])
        assert (one.origin, one.x_scale) == (batch.origin, batch.x_scale)
        assert one.y_scale == batch.y_scale[i]
    assert batch.y_scale[4] == 1.0  # an all-zero row keeps its scale at 1
    q = np.random.default_rng(3).normal(size=(6, 8))
    rows = np.stack([Coordinates.from_data(x, y[i]).physical("two_lorentzian", q[i])
                     for i in range(6)])
    np.testing.assert_array_equal(batch.physical("two_lorentzian", q), rows)
    np.testing.assert_array_equal(batch.y(y)[1], Coordinates.from_data(x, y[1]).y(y[1]))


@pytest.mark.parametrize("x,y", [([], []), ([1, 2], [1]), ([1, np.nan], [1, 2]),
                                 ([1, 2], [1, np.inf]), ([2, 2], [1, 1])])
def test_frame_argument_errors(x, y):
    with pytest.raises(ValueError):
        Coordinates.from_data(x, y)


@pytest.mark.gpu
@pytest.mark.parametrize("model", list(P))
def test_normalized_fit_in_far_units(hip, oracle_port, model):
    """tests/numerics_checks.hpp:188-240: a grid at 1e6 ± 0.005 with signals of
    1e-6. The device fits the normalised float32 arrays; parameters mapped
    back in float64 agree with the float64 oracle fitted in the same frame
    (compensated sums, scaled Cholesky) to 1.5e-4 of the unit parameters."""
    p = P[model]
    n = 16
    base = Coordinates(1e6, .005, 1e-6)
    rng = np.random.default_rng(31)
    unit = np.tile(np.array(BASE[p], np.float64), (n, 1))
    unit[:, 0] *= rng.uniform(0.8, 1.2, n)
    unit[:, 1] += rng.uniform(-0.05, 0.05, n)
    phys = base.physical(model, unit)
    x = base.origin + base.x_scale * (-1 + np.arange(257) / 128)
    y = np.stack([model64(model, x, q) for q in phys])
    lower = np.full(p, -np.inf)
    upper = np.full(p, np.inf)
    lower[0], lower[2] = 0, 1e-5
    if p == 4:
        lower[3], upper[3] = 0, 1
    if p == 8:
        lower[3], lower[5] = 0, 1e-5
    got, frame = fit_normalized(x, y, model, None, (lower, upper), uncertainties=True)
    from oracle import oracle
    ref = oracle_port.fit(frame.x(x), frame.y(y), model, None, (lower, upper), threads=8,
                          options=oracle.make_options(compensated=True, scaled=True))
    assert np.all(ref.status == 0) and np.all(got.status == 0), (got.status, ref.status)
    err = np.abs(frame.normalize(model, got.params) - ref.params)
    assert err.max() <= 1.5e-4, err.max(axis=0)
    assert got.params.dtype == np.float64 and got.covariance.shape == (n, p, p)
    # Fitting the raw float32 arrays instead loses the grid: x rounds to a
    # handful of distinct float32 values around 1e6.
    assert len(np.unique(x.astype(np.float32))) < len(x) // 4
