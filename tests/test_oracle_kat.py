"""Closed-form, special-case and independent-library checks of the oracle
restatement: each pins arithmetic the golden vectors alone leave open, chosen
so that a dropped term, a wrong sign or index, or a transposed operand fails.
Sources are the reference's own checks (tests/*_checks.hpp,
python/tests/*.py) and scipy.optimize.curve_fit."""
import numpy as np
import pytest

import datasets
from oracle import oracle

F32, F64 = np.float32, np.float64


# --- Cholesky and compensated summation (tests/solver_checks.hpp:44-55,
# tests/numerics_checks.hpp:95-106) -----------------------------------------

@pytest.mark.parametrize("dtype", [F32, F64])
def test_cholesky_known_answers(oracle_port, dtype):
    a = np.array([[4, 12, -16], [12, 37, -43], [-16, -43, 98]])
    ok, _, b = oracle_port.cholesky(a, [-20, -43, 192], dtype=dtype)
    assert ok and np.allclose(b, [1, 2, 3], atol=1e-5)
    assert not oracle_port.cholesky([[1, 2], [2, 1]], [1, 2], dtype=dtype)[0]
    assert not oracle_port.cholesky([[1, 1], [1, 1]], [1, 2], dtype=dtype)[0]
    for scaled in (False, True):
        ok, _, b = oracle_port.cholesky([[4, 1, 0], [1, 3, 1], [0, 1, 2]], [6, 10, 8],
                                        scaled=scaled, dtype=dtype)
        assert ok and np.allclose(b, [1, 2, 3], atol=1e-6 if dtype == F32 else 1e-14)


def test_cholesky_matches_numpy_on_random_spd(oracle_port):
    rng = np.random.default_rng(3)
    for n in range(1, 9):
        j = rng.normal(size=(3 * n, n))
        a = j.T @ j + 0.1 * np.eye(n)
        b = rng.normal(size=n)
        for scaled in (False, True):
            ok, _, x = oracle_port.cholesky(a, b, scaled=scaled)
            assert ok
            np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)


def test_kahan_keeps_small_terms():
    lib = oracle.load("port").lib
    import ctypes
    lib.oracle_kahan_sum_f32.restype = ctypes.c_float
    lib.oracle_kahan_sum_f32.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float]
    ones = np.ones(1000, F32)
    assert lib.oracle_kahan_sum_f32(ones.ctypes.data, 1000, 1e8) == np.float32(100001000)


# --- Models: hand values, endpoints, finite differences
# (tests/model_checks.hpp:54-146) -------------------------------------------

@pytest.mark.parametrize("dtype", [F32, F64])
def test_model_hand_values(oracle_port, dtype):
    tol = 2e-6 if dtype == F32 else 1e-14

    def value(model, x, p):
        return float(oracle_port.model(model, x, p, dtype)[0])

    single = [4, 1, 2]
    assert value("gaussian", 1, single) == pytest.approx(4, rel=tol, abs=tol)
    assert value("gaussian", 3, single) == pytest.approx(4 * np.exp(-0.5), rel=tol, abs=tol)
    assert value("lorentzian", 1, single) == pytest.approx(4, rel=tol, abs=tol)
    assert value("lorentzian", 3, single) == pytest.approx(2, rel=tol, abs=tol)
    voigt = [4, 1, 2, 0.25]
    assert value("pseudo_voigt", 1, voigt) == pytest.approx(4, rel=tol, abs=tol)
    assert value("pseudo_voigt", 2, voigt) == pytest.approx(2, rel=tol, abs=tol)
    assert value("pseudo_voigt", 3, voigt) == pytest.approx(0.3875, rel=tol, abs=tol)
    pair = [4, 1, 2, 3, -1, 1, 0.5, -0.25]
    assert value("two_lorentzian", 1, pair) == pytest.approx(4.85, rel=tol, abs=tol)
    assert value("two_lorentzian", -1, pair) == pytest.approx(5.75, rel=tol, abs=tol)
    for x in (-3, 1, 2, 5):
        g = value("gaussian", x, [4, 1, 2 / np.sqrt(8 * np.log(2))])
        l = value("lorentzian", x, [4, 1, 1])
        assert value("pseudo_voigt", x, [4, 1, 2, 0]) == pytest.approx(g, rel=tol, abs=tol)
        assert value("pseudo_voigt", x, [4, 1, 2, 1]) == pytest.approx(l, rel=tol, abs=tol)


@pytest.mark.parametrize("model", list(datasets.MODELS))
@pytest.mark.parametrize("dtype", [F32, F64])
def test_jacobian_finite_differences(oracle_port, model, dtype):
    paired = model == "two_lorentzian"
    p = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8}[model]
    rng = np.random.default_rng(0xBA7CF17)
    step, atol, rtol = (0.002, 3e-4, 2e-3) if dtype == F32 else (1e-5, 2e-8, 2e-7)
    for trial in range(400):
        x = rng.uniform(-6, 6)
        q = np.array([rng.uniform(-5, 5), rng.uniform(-4, 4), rng.uniform(0.3, 3),
                      rng.uniform(-5, 5) if paired else rng.uniform(0, 1), rng.uniform(-4, 4),
                      rng.uniform(0.3, 3), rng.uniform(-2, 2), rng.uniform(-1, 1)])
        if trial % 10 == 0:
            q[0] = 0
            if paired:
                q[3] = 0
        if trial % 11 == 0:
            x = q[1]
        q = q.astype(dtype)
        x = dtype(x)
        _, jac = oracle_port.model(model, x, q, dtype)
        for k in range(p):
            peak = k in (1, 2) or (paired and k in (4, 5))
            scale = q[2 if k < 3 else 5] if peak else max(1, abs(q[k]))
            up, dn = q.copy(), q.copy()
            up[k] += dtype(step * scale)
            dn[k] -= dtype(step * scale)
            numeric = (float(oracle_port.model(model, x, up, dtype)[0])
                       - float(oracle_port.model(model, x, dn, dtype)[0])) / float(up[k] - dn[k])
            assert abs(jac[k] - numeric) <= atol + rtol * abs(numeric), (trial, k)


# --- Solver special cases (tests/solver_checks.hpp:175-346) ------------------

@pytest.mark.parametrize("dtype", [F32, F64])
def test_bounds_and_stopping_rules(oracle_port, dtype):
    fit = oracle_port.fit
    x, y = np.array([-1, 0, 1], dtype), np.array([3, 3, 3], dtype)
    r = fit(x, y, "constant", [0], ([0], [2]))
    assert r.status[0] == 0 and r.params[0, 0] == 2 and abs(r.chi2[0] - 3) < 1e-6
    r = fit(x, y, "constant", [99], ([1], [1]))
    assert r.status[0] == 0 and r.params[0, 0] == 1 and r.iterations[0] == 0
    assert abs(r.chi2[0] - 12) < 1e-6
    r = fit(x, -y, "constant", [0], ([-2], [2]))
    assert r.status[0] == 0 and r.params[0, 0] == -2
    r = fit(np.array([0, 1, 2], dtype), np.array([3, 5, 7], dtype), "affine", [0, 0],
            ([0, -np.inf], [2, np.inf]))
    assert r.status[0] == 0 and r.params[0, 0] == 2
    assert abs(r.params[0, 1] - 2.6) < 2e-5 and abs(r.chi2[0] - 1.2) < 2e-6
    zeros = np.zeros(3, dtype)
    r = fit(zeros, y, "lorentzian", [1, 0, 1], ([0, 0, 1], [10, 0, 1]))
    assert r.status[0] == 0 and abs(r.params[0, 0] - 3) < 1e-5
    px = (-5 + 10 * np.arange(129) / 128).astype(dtype)
    truth = np.array([2, 0.2, 0.7], dtype)
    py = np.array([oracle_port.model("gaussian", v, truth, dtype)[0] for v in px], dtype)
    r = fit(px, py, "gaussian", [1.5, -0.1, -0.5], ([-np.inf, -np.inf, 0.05], [np.inf, np.inf, 2]))
    assert r.status[0] == 0 and r.params[0, 2] > 0
    np.testing.assert_allclose(r.params[0], truth, atol=1e-5)
    o = oracle.make_options(gradient_tolerance=10)
    r = fit(x, y, "constant", [0], None, options=o)
    assert r.status[0] == 0 and r.iterations[0] == 0 and r.params[0, 0] == 0
    o = oracle.make_options(gradient_tolerance=0, cost_tolerance=0, step_tolerance=10)
    r = fit(x, y, "constant", [0], None, options=o)
    assert r.status[0] == 0 and r.iterations[0] == 1
    o = oracle.make_options(gradient_tolerance=0, step_tolerance=0, cost_tolerance=1)
    r = fit(x, y, "constant", [0], None, options=o)
    assert r.status[0] == 0 and r.iterations[0] == 1
    o = oracle.make_options(max_iterations=3, initial_lambda=16 / np.finfo(dtype).eps)
    r = fit(x, y, "constant", [0], None, options=o)
    assert r.status[0] == 1 and r.iterations[0] == 3


@pytest.mark.parametrize("dtype", [F32, F64])
def test_failure_statuses(oracle_port, dtype):
    fit = oracle_port.fit
    x, y = np.ones(3, dtype), np.full(3, 2, dtype)
    r = fit(x, y, "constant", [0], None, options=oracle.make_options(max_iterations=0))
    assert r.status[0] == 1 and r.iterations[0] == 0 and r.chi2[0] == 12
    ey = np.full(3, np.exp(2.0), dtype)
    r = fit(x, ey, "exponential", [-2], None, options=oracle.make_options(max_iterations=1))
    assert r.status[0] == 1 and r.iterations[0] == 1 and r.params[0, 0] == -2
    r = fit(x, ey, "exponential", [-2], None, options=oracle.make_options(max_iterations=200))
    assert r.status[0] == 0 and r.iterations[0] > 1 and abs(r.params[0, 0] - 2) < 1e-5
    r = fit(np.zeros(3, dtype), y, "lorentzian", [1, 0, 1], None)
    assert r.status[0] == 2 and r.iterations[0] == 1 and r.chi2[0] == 3
    assert list(r.params[0]) == [1, 0, 1]
    bad_cases = [
        dict(x=np.array([np.nan, 1, 1], dtype)), dict(y=np.array([np.inf, 2, 2], dtype)),
        dict(p0=[np.nan]), dict(options=oracle.make_options(initial_lambda=0)),
        dict(options=oracle.make_options(gradient_tolerance=-1)),
        dict(options=oracle.make_options(step_tolerance=np.inf)),
        dict(options=oracle.make_options(cost_tolerance=np.nan)),
        dict(bounds=([2], [1])), dict(bounds=([np.nan], [np.inf])),
        dict(bounds=([np.inf], [np.inf])), dict(bounds=([-np.inf], [-np.inf])),
    ]
    for case in bad_cases:
        r = fit(case.get("x", x), case.get("y", y), "constant", case.get("p0", [0]),
                case.get("bounds"), options=case.get("options"))
        assert r.status[0] == 3 and r.iterations[0] == 0 and np.isnan(r.chi2[0]), case
    r = fit(np.zeros(3, dtype), y, "gaussian", [1, 0, 0], None)
    assert r.status[0] == 3


def test_recovery_invariants_and_scipy(oracle_port):
    """Independent oracle: SciPy's bounded TRF optimum on the reference's SciPy
    datasets (python/tests/test_scipy.py:8-23), float64 tolerance 2e-7 on
    parameters; plus accepted-objective and feasibility invariants."""
    from scipy.optimize import curve_fit
    for model, f in datasets.MODELS.items():
        x, y, p0, bounds, _ = datasets.comparison_set(model, 60, 129, seed=20261109)
        r = oracle_port.fit(x, y, model, p0, bounds, threads=4)
        assert np.all(r.status == 0)
        lo, hi = bounds
        assert np.all((r.params >= lo) & (r.params <= hi))
        c0 = np.array([np.sum((row - f(x, *p0)) ** 2) for row in y])
        assert np.all(r.chi2 <= c0)
        recomputed = np.array([np.sum((row - f(x, *p)) ** 2) for row, p in zip(y, r.params)])
        np.testing.assert_allclose(r.chi2, recomputed, rtol=1e-12, atol=1e-12)
        for i in range(0, 60, 6):
            ref = curve_fit(f, x, y[i], p0=p0, bounds=bounds, max_nfev=1000, ftol=1e-11,
                            xtol=1e-11, gtol=1e-11)[0]
            np.testing.assert_allclose(r.params[i], ref, rtol=2e-7, atol=2e-7)


# --- Covariance (tests/uncertainty_checks.hpp:20-72,
# python/tests/test_uncertainty.py) ------------------------------------------

@pytest.mark.parametrize("dtype", [F32, F64])
def test_covariance_algebra(oracle_port, dtype):
    x = np.array([-2, -1, 0, 1, 2], dtype)
    unc = oracle_port.uncertainty
    cov, se, st = unc("line", x, [1, 2], 0, 6, dtype=dtype)
    assert st == 0
    assert cov[0, 0] == pytest.approx(.4, rel=1e-6, abs=1e-7)
    assert cov[1, 1] == pytest.approx(.2, rel=1e-6, abs=1e-7)
    assert cov[0, 1] == 0 and cov[1, 0] == 0
    cov, se, st = unc("line", x, [1, 2], 0, 6, ([1, -np.inf], [1, np.inf]), dtype=dtype)
    assert st == 0 and se[0] == 0 and cov[1, 1] == pytest.approx(.15, rel=1e-6, abs=1e-7)
    cov, se, st = unc("line", x, [1, 2], 0, 6, ([1, 2], [1, 2]), dtype=dtype)
    assert st == 0 and se[1] == 0
    cov, se, st = unc("line", x[:2], [1, 2], 0, 6, dtype=dtype)
    assert st == 2 and np.isnan(cov[0, 0])
    cov, se, st = unc("duplicate", x, [1, 2], 0, 6, dtype=dtype)
    assert st == 3 and np.isnan(se[0])
    assert unc("line", x, [1, 2], 0, 6, ([1, -np.inf], [np.inf, np.inf]), dtype=dtype)[2] == 4
    assert unc("line", x, [1, 2], 1, 6, dtype=dtype)[2] == 1
    assert unc("line", x, [1, 2], 0, np.nan, dtype=dtype)[2] == 5
    cov, se, st = unc("line", x, [1, 2], 0, 0, dtype=dtype)
    assert st == 0 and se[0] == 0


@pytest.mark.parametrize("model", list(datasets.MODELS))
def test_covariance_matches_scipy(oracle_port, model):
    from scipy.optimize import curve_fit
    f = datasets.MODELS[model]
    x, y, p0, bounds, _ = datasets.comparison_set(model, 8, 129, seed=20261109)
    r = oracle_port.fit(x, y, model, p0, bounds, uncertainties=True)
    assert np.all(r.uncertainty_status == 0)
    for i in range(8):
        _, expected = curve_fit(f, x, y[i], p0=p0, bounds=bounds, max_nfev=1000, ftol=1e-11,
                                xtol=1e-11, gtol=1e-11)
        np.testing.assert_allclose(r.covariance[i], expected, rtol=2e-5, atol=1e-10)
    np.testing.assert_array_equal(r.covariance, r.covariance.transpose(0, 2, 1))
    np.testing.assert_allclose(r.stderr ** 2, np.diagonal(r.covariance, axis1=1, axis2=2),
                               rtol=2e-7, atol=1e-16)


TRUTH = {"gaussian": [2, .2, .7], "lorentzian": [2, .2, .7], "pseudo_voigt": [2, .2, 1.4, .5],
         "two_lorentzian": [2, -1.5, .5, 1.7, 1.5, .6, .25, .01]}
INITIAL = {"gaussian": [1.8, 0, .8], "lorentzian": [1.8, 0, .8], "pseudo_voigt": [1.8, 0, 1.6, .45],
           "two_lorentzian": [1.8, -1.3, .55, 1.5, 1.3, .7, .2, 0]}


def domain(p):
    lower, upper = np.full(p, -np.inf), np.full(p, np.inf)
    lower[0], lower[2] = 0, 1e-4
    lower[1], upper[1] = -5, 5
    if p == 4:
        lower[3], upper[3] = 0, 1
    if p == 8:
        lower[3], lower[5], lower[4], upper[4] = 0, 1e-4, -5, 5
    return lower, upper


@pytest.mark.parametrize("model", list(datasets.MODELS))
def test_monte_carlo_coverage(oracle_port, model):
    """python/tests/test_uncertainty.py:101-118: 10,000 realisations, the
    ±1σ intervals cover the truth 68.27% ± 2 pp for every parameter."""
    n = 10000
    rng = np.random.default_rng(20261110 + list(datasets.MODELS).index(model))
    truth = np.asarray(TRUTH[model])
    x = np.linspace(-5, 5, 129)
    y = datasets.MODELS[model](x, *truth) + rng.normal(0, .03, (n, x.size))
    r = oracle_port.fit(x, y, model, INITIAL[model], domain(truth.size), uncertainties=True,
                        threads=8)
    ok = (r.status == 0) & (r.uncertainty_status == 0)
    assert np.all(ok)
    coverage = np.count_nonzero(np.abs(r.params - truth) <= r.stderr, axis=0) / n
    assert np.all(np.abs(coverage - 0.682689492137) <= .02), coverage


# --- Initial guess (tests/initial_guess_checks.hpp:22-176) -------------------

@pytest.mark.parametrize("dtype", [F32, F64])
def test_guess_features(oracle_port, dtype):
    tol = 2e-6 if dtype == F32 else 1e-12
    x = np.arange(9, dtype=dtype)
    y = np.array([2, 2, 2, 4, 6, 4, 2, 2, 2], dtype)
    g, b, s = oracle_port.guess(x, y, "gaussian")
    assert s[0] == 0 and abs(b[0] - 2) <= tol
    assert abs(g[0, 0] - 4) <= tol and abs(g[0, 1] - 4) <= tol
    assert abs(g[0, 2] - 2 / 2.354820045030949382) <= tol
    g, _, s = oracle_port.guess(x, y, "lorentzian")
    assert s[0] == 0 and abs(g[0, 2] - 1) <= tol
    g, _, s = oracle_port.guess(x, y, "pseudo_voigt")
    assert s[0] == 0 and abs(g[0, 2] - 2) <= tol and abs(g[0, 3] - .5) <= tol
    y2 = np.array([1, 1, 1, 4, 5, 3.5, 1, 1, 1], dtype)
    g, _, s = oracle_port.guess(x, y2, "lorentzian")
    assert s[0] == 0 and abs(g[0, 1] - 37.5 / 9.5) <= tol
    assert abs(g[0, 2] - (5.2 - 8.0 / 3) / 2) <= tol
    xn = np.array([10, 11, 12, 14, 15, 16, 18, 19, 20], dtype)
    g, _, s = oracle_port.guess(xn, y, "lorentzian")
    assert s[0] == 0 and abs(g[0, 1] - 15) <= tol and abs(g[0, 2] - 1) <= tol


@pytest.mark.parametrize("dtype", [F32, F64])
def test_guess_paired_and_invalid(oracle_port, dtype):
    x = (-5 + 10 * np.arange(201) / 200).astype(dtype)
    for reverse in (False, True):
        truth = np.array([2 if reverse else 3, -1.5, .4, 3 if reverse else 2, 1.5, .5, .3, -.04])
        y = datasets.two_lorentzian(x.astype(np.float64), *truth).astype(dtype)
        g, _, s = oracle_port.guess(x, y, "two_lorentzian")
        assert s[0] == 0 and g[0, 1] < g[0, 4]
        for k in range(6):
            assert abs(g[0, k] - truth[k]) <= (.15 if k % 3 == 0 else .06)
        assert abs(g[0, 6] - truth[6]) <= .12 and abs(g[0, 7] - truth[7]) <= .02
    y = (0.2 + datasets.lorentzian(x.astype(np.float64), 2, 0, .5)).astype(dtype)
    assert oracle_port.guess(x, y, "two_lorentzian")[2][0] == 3
    xs = np.arange(9, dtype=dtype)
    zeros = np.zeros(9, dtype)
    g, _, s = oracle_port.guess(xs, zeros, "gaussian")
    assert s[0] == 2 and np.all(np.isnan(g))
    for bad_x, bad_y in ((np.r_[0, 1, 1, 3:9].astype(dtype), zeros),
                         (np.r_[0, 1, -1, 3:9].astype(dtype), zeros),
                         (np.r_[0, 1, np.inf, 3:9].astype(dtype), zeros),
                         (xs, np.r_[0, 0, np.nan, np.zeros(6)].astype(dtype))):
        assert oracle_port.guess(bad_x, bad_y, "gaussian")[2][0] == 1
    assert oracle_port.guess(xs[:4], zeros[:4], "gaussian")[2][0] == 1
    clipped = datasets.gaussian(xs.astype(np.float64), 2, 0, 1).astype(dtype)
    assert oracle_port.guess(xs, clipped, "gaussian")[2][0] == 2
    dip = datasets.gaussian(xs.astype(np.float64), -2, 4, 1).astype(dtype)
    assert oracle_port.guess(xs, dip, "gaussian")[2][0] == 2
