"""Pin the oracle restatement (oracle/bf_oracle.hpp) to the golden vectors that
the reference's own CPU headers wrote (tests/golden/make_golden.py over
oracle/_ref). Both are built with -ffp-contract=off, so agreement is bitwise."""
import numpy as np
import pytest

import datasets

FIELDS = ("params", "status", "iterations", "chi2")


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint8)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


def digest(*arrays):
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return np.frombuffer(h.digest(), np.uint8)


@pytest.mark.parametrize("name,dtype", [("f32", np.float32), ("f64", np.float64)])
def test_config1_bitwise(oracle_port, golden, name, dtype):
    g = golden["config1"]
    x, y, bounds = datasets.config1(dtype)
    assert np.array_equal(digest(x, y), g[f"c1_{name}_digest"]), "input generator drifted"
    r = oracle_port.fit(x, y, "lorentzian", None, bounds, threads=4)
    for field in FIELDS:
        assert same(getattr(r, field), g[f"c1_{name}_{field}"]), field
    params, baseline, status = oracle_port.guess(x, y, "lorentzian")
    assert same(params, g[f"c1_{name}_guess"])
    assert same(baseline, g[f"c1_{name}_baseline"])
    assert same(status, g[f"c1_{name}_gstatus"])
    assert np.all(r.status == 0) and r.iterations.max() <= 100


@pytest.mark.parametrize("model", list(datasets.MODELS))
@pytest.mark.parametrize("m,n", [(32, 257), (129, 257), (1024, 65)])
@pytest.mark.parametrize("name,dtype", [("f32", np.float32), ("f64", np.float64)])
def test_sets_bitwise(oracle_port, golden, model, m, n, name, dtype):
    from oracle import oracle
    g = golden["sets"]
    x, y, p0, bounds, _ = datasets.comparison_set(model, n, m)
    key = f"{model}_{m}"
    assert np.array_equal(digest(x, y, p0), g[f"{key}_digest"]), "input generator drifted"
    xs, ys = x.astype(dtype), y.astype(dtype)
    given = oracle_port.fit(xs, ys, model, p0, bounds, uncertainties=True, threads=4)
    guessed = oracle_port.fit(xs, ys, model, None, bounds, threads=4)
    comp = oracle_port.fit(xs, ys, model, p0, bounds, threads=4,
                           options=oracle.make_options(compensated=True, scaled=True))
    for tag, r in (("given", given), ("guessed", guessed), ("comp", comp)):
        for field in FIELDS:
            assert same(getattr(r, field), g[f"{key}_{name}_{tag}_{field}"]), (tag, field)
    assert same(given.covariance, g[f"{key}_{name}_given_cov"])
    assert same(given.stderr, g[f"{key}_{name}_given_stderr"])
    assert same(given.uncertainty_status, g[f"{key}_{name}_given_ustatus"])
    params, baseline, status = oracle_port.guess(xs, ys, model)
    assert same(params, g[f"{key}_{name}_guess"])
    assert same(baseline, g[f"{key}_{name}_baseline"])
    assert same(status, g[f"{key}_{name}_gstatus"])


def test_failure_fixtures_bitwise(oracle_port, golden):
    from oracle import oracle
    g = golden["failures"]
    x, y, p0 = datasets.isolation_fixture()
    for flags in range(4):
        opts = oracle.make_options(compensated=flags & 1, scaled=flags & 2)
        r = oracle_port.fit(x, y, "lorentzian", p0, None, options=opts)
        for field in FIELDS:
            assert same(getattr(r, field), g[f"iso_{flags}_{field}"])
    # The reference's documented expectation (tests/metal_lm_checks.hpp:67-72).
    assert list(g["iso_0_status"]) == [2, 0, 3] and list(g["iso_0_iterations"]) == [1, 0, 0]
    assert g["iso_0_chi2"][0] == 3
    for model in datasets.MODELS:
        xs, inits, _ = datasets.pathology_fixture(model)
        ys = g[f"path_{model}_y"]
        for flags in range(4):
            opts = oracle.make_options(compensated=flags & 1, scaled=flags & 2)
            r = oracle_port.fit(xs, ys, model, inits, None, options=opts)
            for field in FIELDS:
                assert same(getattr(r, field), g[f"path_{model}_{flags}_{field}"]), (model, field)
            # tests/numerics_checks.hpp:320-332 of the reference.
            assert list(r.status) == [4, 4, 3, 3, 3, 0, 5, 0, 3]
    for name, dtype in (("f32", np.float32), ("f64", np.float64)):
        x, y, p0 = datasets.api_isolation_fixture(dtype)
        r = oracle_port.fit(x, y, "lorentzian", p0, None)
        assert same(r.status, g[f"api_{name}_status"]) and same(r.params, g[f"api_{name}_params"])
        assert list(r.status) == [0, 3, 3, 4, 0]


def test_batch_equals_single_rows(oracle_port):
    """tests/batch_checks.hpp:142-158: a batch row is bitwise the single fit,
    whatever the worker count."""
    x, y, p0, bounds, _ = datasets.comparison_set("pseudo_voigt", 130, 129)
    one = oracle_port.fit(x, y, "pseudo_voigt", p0, bounds, threads=1)
    many = oracle_port.fit(x, y, "pseudo_voigt", p0, bounds, threads=8)
    for field in FIELDS:
        assert same(getattr(one, field), getattr(many, field))
    single = oracle_port.fit(x, y[57:58], "pseudo_voigt", p0, bounds)
    for field in FIELDS:
        assert same(getattr(single, field)[0], getattr(one, field)[57])
