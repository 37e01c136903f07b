"""GPU parity of the HIP path against the CPU oracle on identical inputs.

Tolerances (written here, per test):
  * vs the FP32 oracle: the reference's own GPU gate (tests/metal_lm_checks.hpp:41-45)
    params |gpu−cpu| ≤ 2e-4 + 2e-4·|cpu|, χ² ≤ 1e-6 + 2e-5·|cpu|, equal status.
  * vs the FP64 oracle on the metric workload: status agreement ≥ 99.9 %;
    converged rows amplitude and width relative error ≤ 1e-4, centre error
    ≤ 1e-4·width, χ² relative error ≤ 1e-5.
  * the data-only guess is integer/branch work on FP32 values and must equal
    the FP32 oracle bit for bit.
"""
import ctypes

import numpy as np
import pytest

import datasets

pytestmark = pytest.mark.gpu

MAPPINGS = ["lane", "group4_streamed", "group8", "group16", "group32", "group64", "group256"]
NUMERICS = [0, 5]  # FP32 as the reference GPU kernels; compensated FP32 + FP64 solve
P = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8}


def mappings_for(m):
    """Mappings whose register tile (at most 32 points per lane) holds M, and
    the 4-lane groups, which always stream."""
    return [mp for mp in MAPPINGS
            if mp in ("lane", "group4_streamed", "group64", "group256")
            or (mp == "group32" and m <= 1024)
            or (mp == "group16" and m <= 512) or (mp == "group8" and m <= 256)]


def assert_fp32_gate(got, ref, label):
    bad = np.nonzero(got.status != ref.status)[0]
    assert bad.size == 0, f"{label}: status differs at rows {bad[:10]} " \
                          f"gpu={got.status[bad[:10]]} cpu={ref.status[bad[:10]]}"
    ok = ref.status == 0
    d = np.abs(got.params[ok] - ref.params[ok]) - (2e-4 + 2e-4 * np.abs(ref.params[ok]))
    assert np.all(d <= 0), f"{label}: params off by {d.max():.3g} beyond tolerance"
    dc = np.abs(got.chi2[ok] - ref.chi2[ok]) - (1e-6 + 2e-5 * np.abs(ref.chi2[ok]))
    assert np.all(dc <= 0), f"{label}: chi2 off by {dc.max():.3g} beyond tolerance"
    assert np.all(got.iterations <= 100)


def test_model_values_and_jacobians(hip, oracle_port):
    """Device FP32 model math vs the float64 oracle (tests/models_metal.mm:19-60)."""
    from batchfit_amd import _native
    lib = _native.library()
    rng = np.random.default_rng(17)
    for model, p in P.items():
        n = 2048
        x = rng.uniform(-6, 6, n).astype(np.float32)
        q = np.zeros((n, 8), np.float32)
        q[:, 0] = rng.uniform(-5, 5, n)
        q[:, 1] = rng.uniform(-4, 4, n)
        q[:, 2] = rng.uniform(0.3, 3, n)
        q[:, 3] = rng.uniform(-5, 5, n) if model == "two_lorentzian" else rng.uniform(0, 1, n)
        q[:, 4] = rng.uniform(-4, 4, n)
        q[:, 5] = rng.uniform(0.3, 3, n)
        q[:, 6] = rng.uniform(-2, 2, n)
        q[:, 7] = rng.uniform(-1, 1, n)
        value = np.empty(n, np.float32)
        jac = np.empty((n, p), np.float32)
        err = _native.error_buffer()
        _native.check(lib.bf_hip_model_eval(_native.MODELS[model], n, x.ctypes.data,
                                            q.ctypes.data, value.ctypes.data, jac.ctypes.data,
                                            err, len(err)), err)
        for i in range(n):
            rv, rj = oracle_port.model(model, float(x[i]), q[i].astype(np.float64))
            assert abs(value[i] - rv) <= 4e-6 * (1 + abs(rv)), (model, i, value[i], rv)
            assert np.all(np.abs(jac[i] - rj) <= 1e-5 * (1 + np.abs(rj))), (model, i)


def test_guess_bitwise_config1(hip, oracle_port):
    from batchfit_amd import _native
    x, y, _ = datasets.config1(np.float32)
    got = device_guess(x, y, "lorentzian")
    ref = oracle_port.guess(x, y, "lorentzian")
    for a, b in zip(got, ref):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                              np.ascontiguousarray(b).view(np.uint8))


GUESS_AUTO, GUESS_WAVE, GUESS_WIDE = 0, 1, 2  # bf_guess_form


def device_guess(x, y, model, form=GUESS_AUTO):
    """bf_hip_guess (form GUESS_AUTO) or bf_hip_guess_form with a chosen kernel:
    GUESS_WAVE stages each fit's signal in LDS, GUESS_WIDE reads it from memory."""
    from batchfit_amd import _native
    lib = _native.library()
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    n, m = y.shape
    params = np.empty((n, P[model]), np.float32)
    baseline = np.empty(n, np.float32)
    status = np.empty(n, np.int32)
    err = _native.error_buffer()
    if form == GUESS_AUTO:
        rc = lib.bf_hip_guess(_native.MODELS[model], n, m, x.ctypes.data, y.ctypes.data,
                              params.ctypes.data, baseline.ctypes.data, status.ctypes.data,
                              err, len(err))
    else:
        rc = lib.bf_hip_guess_form(_native.MODELS[model], form, n, m, x.ctypes.data,
                                   y.ctypes.data, params.ctypes.data, baseline.ctypes.data,
                                   status.ctypes.data, err, len(err))
    _native.check(rc, err)
    return params, baseline, status


def bitwise(got, ref):
    return all(np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                              np.ascontiguousarray(b).view(np.uint8)) for a, b in zip(got, ref))


@pytest.mark.parametrize("model", list(P))
@pytest.mark.parametrize("m,n", [(32, 257), (129, 257), (1024, 65)])
def test_guess_bitwise_sets(hip, oracle_port, model, m, n):
    x, y, _, _, _ = datasets.comparison_set(model, n, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    got = device_guess(x, y, model)
    ref = oracle_port.guess(x, y, model)
    for a, b in zip(got, ref):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                              np.ascontiguousarray(b).view(np.uint8)), model


@pytest.mark.parametrize("model", list(P))
@pytest.mark.parametrize("m", [1500, 2600, 4096, 6000])
def test_guess_bitwise_long_windows(hip, oracle_port, model, m):
    """Edge windows of 150–600 points: medians by register selection (up to
    1024 values) and by value search beyond. Bitwise on noisy data; on a
    quantised copy, where many equal values (zeros of both signs among them)
    meet at the median, equal in value: which of two equal elements the
    reference's nth_element leaves in the middle is unspecified, so the sign
    of a zero median is not part of the contract."""
    x, y, _, _, _ = datasets.comparison_set(model, 48, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    got = device_guess(x, y, model)
    ref = oracle_port.guess(x, y, model, threads=8)
    for a, b in zip(got, ref):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                              np.ascontiguousarray(b).view(np.uint8)), (model, m)
    q = (np.round(y * 50) / 50).astype(np.float32)
    got = device_guess(x, q, model)
    ref = oracle_port.guess(x, q, model, threads=8)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b, equal_nan=True), (model, m, "quantised")


def prominence_rows(m, n, rng):
    """Doublet rows that make the second-peak search walk far: one peak only,
    pure noise, and second peaks from well below to well above three noise
    scales, on absolute noise."""
    x = np.linspace(-10, 10, m).astype(np.float32)
    a1 = rng.uniform(0.5, 5, n)
    c1 = rng.uniform(-8, 8, n)
    w1 = rng.uniform(0.05, 1.5, n)
    kind = np.arange(n) % 4
    a2 = np.where(kind == 0, 0.0, np.where(kind == 1, 0.0, rng.uniform(0.005, 0.2, n)))
    a1 = np.where(kind == 1, 0.0, a1)  # pure noise rows
    c2 = rng.uniform(-9, 9, n)
    w2 = rng.uniform(0.05, 1.0, n)

    def lor(a, c, w):
        return a[:, None] / (1 + ((x[None, :] - c[:, None]) / w[:, None]) ** 2)

    y = lor(a1, c1, w1) + lor(a2, c2, w2) + rng.normal(0, 0.01, (n, m))
    return x, y.astype(np.float32)


@pytest.mark.parametrize("m", [700, 2048, 4096])
def test_two_peak_guess_prominence_walks(hip, oracle_port, m):
    """The two-peak guess's search for a second prominent maximum on rows that
    make it walk far. The wave tests each candidate 64 points of a side at a
    time; the result must equal the oracle's point-by-point walk bit for bit,
    with the signal staged in LDS and read from memory alike."""
    x, y = prominence_rows(m, 96, np.random.default_rng(m))
    ref = oracle_port.guess(x, y, "two_lorentzian", threads=8)
    assert bitwise(device_guess(x, y, "two_lorentzian"), ref), m
    assert bitwise(device_guess(x, y, "two_lorentzian", GUESS_WIDE), ref), (m, "wide")
    # Both outcomes of the search occur in the set.
    assert (ref[2] == 0).any() and (ref[2] == 3).any()


@pytest.mark.parametrize("model", list(P))
@pytest.mark.parametrize("m", [5, 64, 257, 1028, 2600, 6000])
def test_guess_wide_equals_wave_and_oracle(hip, oracle_port, model, m):
    """The wide guess (one wave per fit, the signal formed from y and x at
    every read, edge medians searched in memory, coarse suffix minima) against
    the one-wave guess with the signal staged in LDS and against the oracle,
    bit for bit, on the comparison set and the peak-scan corner rows
    (non-finite observations, an overflowing signal, levels of −0, flat
    tops); on a quantised copy equal in value (the sign of a zero median is
    not part of the contract)."""
    x, y, _, _, _ = datasets.comparison_set(model, 40, m)
    xc, yc = peak_scan_rows(m, np.random.default_rng(m))
    assert np.array_equal(x.astype(np.float32), xc)
    rows = np.vstack([y.astype(np.float32), yc])
    ref = oracle_port.guess(xc, rows, model, threads=8)
    wide = device_guess(xc, rows, model, GUESS_WIDE)
    assert bitwise(wide, ref), (model, m)
    assert bitwise(device_guess(xc, rows, model, GUESS_WAVE), ref), (model, m, "wave")
    with np.errstate(over="ignore"):  # the overflow rows quantise to ±inf
        q = (np.round(rows * 50) / 50).astype(np.float32)
    ref = oracle_port.guess(xc, q, model, threads=8)
    for form in (GUESS_WIDE, GUESS_WAVE):
        for a, b in zip(device_guess(xc, q, model, form), ref):
            assert np.array_equal(a, b, equal_nan=True), (model, m, form, "quantised")


@pytest.mark.parametrize("m", [12289, 16384, 70000])
def test_two_peak_guess_long_spectra(hip, oracle_port, m):
    """Long doublets take the wide guess (the automatic choice beyond 5,120
    points); up to 15,886 points the one-wave guess also holds them, with
    its edge levels searched by value over the staged signal. Both bit for
    bit the oracle's guess (initial_guess.hpp keeps its scratch on the heap,
    :143), at 70,000 points with step minima over runs of four 64-point
    steps; fits from it within the FP32 gate."""
    x, y, _, bounds, _ = datasets.comparison_set("two_lorentzian", 6, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    xw, yw = prominence_rows(m, 12, np.random.default_rng(m))
    for xs, ys in ((x, y), (xw, yw)):
        ref = oracle_port.guess(xs, ys, "two_lorentzian", threads=8)
        assert bitwise(device_guess(xs, ys, "two_lorentzian"), ref), m
        if m <= 15886:
            assert bitwise(device_guess(xs, ys, "two_lorentzian", GUESS_WAVE), ref), (m, "wave")
    assert np.all(oracle_port.guess(x, y, "two_lorentzian", threads=8)[2] == 0)
    fit = hip.fit(x, y, "two_lorentzian", None, bounds, "hip")
    assert_fp32_gate(fit, oracle_port.fit(x, y, "two_lorentzian", None, bounds, threads=8),
                     f"two_lorentzian M={m} guessed")


def test_guess_bitwise_metric_and_doublets(hip, oracle_port):
    x, y, _, _ = datasets.config2(n=1 << 20, rows=4096)
    got = device_guess(x, y, "pseudo_voigt")
    ref = oracle_port.guess(x, y, "pseudo_voigt", threads=8)
    for a, b in zip(got, ref):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                              np.ascontiguousarray(b).view(np.uint8))
    x3, y3, _, _, _, _ = datasets.config3(n=262144, rows=1024)
    got = device_guess(x3, y3, "two_lorentzian")
    ref = oracle_port.guess(x3, y3, "two_lorentzian", threads=8)
    for a, b in zip(got, ref):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                              np.ascontiguousarray(b).view(np.uint8))
    assert bitwise(device_guess(x3, y3, "two_lorentzian", GUESS_WIDE), ref)


@pytest.mark.parametrize("model", list(P))
@pytest.mark.parametrize("m,n", [(32, 257), (129, 257), (1024, 65)])
def test_fit_matches_fp32_oracle(hip, oracle_port, model, m, n):
    x, y, p0, bounds, _ = datasets.comparison_set(model, n, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    ref = oracle_port.fit(x, y, model, p0, bounds, threads=8)
    for mapping in mappings_for(m):
        for numerics in NUMERICS:
            got = hip.fit(x, y, model, p0, bounds, "hip", mapping=mapping, numerics=numerics)
            assert got.params.dtype == np.float32 and got.status.dtype == np.int32
            assert got.iterations.dtype == np.uint32
            assert_fp32_gate(got, ref, f"{model} M={m} {mapping} numerics={numerics}")


def fields_bitwise_equal(a, b, label):
    for name, u, v in zip(a._fields, a, b):
        assert np.array_equal(np.ascontiguousarray(u).view(np.uint8),
                              np.ascontiguousarray(v).view(np.uint8)), f"{label}: {name}"


@pytest.mark.parametrize("model", list(P))
@pytest.mark.parametrize("g,m", [(8, 37), (8, 200), (16, 37), (16, 100), (16, 257), (16, 512), (32, 300),
                                 (32, 1000), (64, 129), (64, 1000), (64, 1500), (64, 2048)])
def test_streamed_equals_register_tile(hip, model, g, m):
    """Points read from memory on every pass give the register tile's result
    bit for bit at the same (lanes, M), partial blocks and empty lanes
    included, with and without uncertainties."""
    x, y, p0, bounds, _ = datasets.comparison_set(model, 97, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    for unc in (False, True):
        a = hip.fit(x, y, model, p0, bounds, "hip", mapping=f"group{g}", uncertainties=unc)
        b = hip.fit(x, y, model, p0, bounds, "hip", mapping=f"group{g}_streamed",
                    uncertainties=unc)
        fields_bitwise_equal(a, b, f"{model} G={g} M={m} uncertainties={unc}")


@pytest.mark.parametrize("n", [1, 2, 7, 8, 9, 15, 16, 17, 33, 100])
def test_streamed_pool_batch_sizes(hip, n):
    """Batches that end inside a wave's pool or a group: the retiring kernel
    with and without the flat-observation test (given p0 / device guess)
    gives the register tile's bytes."""
    for model, m in (("pseudo_voigt", 200), ("two_lorentzian", 100)):
        x, y, p0, bounds, _ = datasets.comparison_set(model, max(n, 8), m)
        x, y = x.astype(np.float32), y[:n].astype(np.float32)
        for start in (p0, None):
            a = hip.fit(x, y, model, start, bounds, "hip", mapping="group8")
            b = hip.fit(x, y, model, start, bounds, "hip", mapping="group8_streamed")
            fields_bitwise_equal(a, b, f"{model} M={m} N={n} p0={'given' if start is not None else 'guess'}")


@pytest.mark.parametrize("model,m", [("gaussian", 256), ("pseudo_voigt", 200),
                                     ("two_lorentzian", 160)])
def test_streamed_equals_register_tile_large_batch(hip, model, m):
    """A batch large enough that every wave of the streamed kernel refills
    its pool many times (fits retire one by one): bit for bit the register
    tile's result, from the device guess, so a skipped or repeated fit shows."""
    x, y, p0, bounds, _ = datasets.comparison_set(model, 40000, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    a = hip.fit(x, y, model, None, bounds, "hip", mapping="group8")
    b = hip.fit(x, y, model, None, bounds, "hip", mapping="group8_streamed")
    fields_bitwise_equal(a, b, f"{model} M={m} N=40000")
    c = hip.fit(x, y, model, None, bounds, "hip", mapping="group64_streamed")
    d = hip.fit(x, y, model, None, bounds, "hip", mapping="group64")
    fields_bitwise_equal(c, d, f"{model} M={m} N=40000, 64 lanes")


@pytest.mark.parametrize("model", list(P))
@pytest.mark.parametrize("g,m", [(8, 700), (16, 1500), (32, 2600)])
def test_streamed_groups_beyond_the_tile(hip, oracle_port, model, g, m):
    """8-, 16- and 32-lane groups stream the points where no register tile
    holds them (more than 32 points per lane)."""
    x, y, p0, bounds, _ = datasets.comparison_set(model, 65, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    ref = oracle_port.fit(x, y, model, p0, bounds, threads=8)
    got = hip.fit(x, y, model, p0, bounds, "hip", mapping=f"group{g}")
    assert_fp32_gate(got, ref, f"{model} M={m} group{g} (streamed)")


@pytest.mark.parametrize("model", list(P))
def test_fit_from_device_guess(hip, oracle_port, model):
    x, y, _, bounds, _ = datasets.comparison_set(model, 257, 129)
    x, y = x.astype(np.float32), y.astype(np.float32)
    ref = oracle_port.fit(x, y, model, None, bounds, threads=8)
    for numerics in NUMERICS:
        got = hip.fit(x, y, model, None, bounds, "hip", numerics=numerics)
        assert_fp32_gate(got, ref, f"{model} guessed numerics={numerics}")


def test_metric_workload_vs_fp64_oracle(hip, oracle_port):
    """16,384 rows of the pseudo-Voigt M=1024 workload, production numerics,
    against the float64 oracle on the same float32 inputs."""
    x, y, bounds, _ = datasets.config2(rows=16384)
    ref = oracle_port.fit(x.astype(np.float64), y.astype(np.float64), "pseudo_voigt", None,
                          bounds, threads=8)
    got = hip.fit(x, y, "pseudo_voigt", None, bounds, "hip")
    agree = got.status == ref.status
    rate = agree.mean()
    for i in np.nonzero(~agree)[0][:20]:
        print(f"status disagreement row {i}: gpu={got.status[i]} fp64={ref.status[i]} "
              f"iterations {got.iterations[i]}/{ref.iterations[i]} params {got.params[i]} "
              f"vs {ref.params[i]}")
    assert rate >= 0.999, rate
    both = (got.status == 0) & (ref.status == 0)
    g, r = got.params[both].astype(np.float64), ref.params[both]
    amp = np.abs(g[:, 0] - r[:, 0]) / np.abs(r[:, 0])
    width = np.abs(g[:, 2] - r[:, 2]) / np.abs(r[:, 2])
    center = np.abs(g[:, 1] - r[:, 1]) / np.abs(r[:, 2])
    chi = np.abs(got.chi2[both] - ref.chi2[both]) / ref.chi2[both]
    print(f"FP64 errors over {both.sum()} converged rows: amplitude max {amp.max():.3g} "
          f"rms {np.sqrt(np.mean(amp**2)):.3g}; width max {width.max():.3g}; centre/width max "
          f"{center.max():.3g}; chi2 max {chi.max():.3g}; status agreement {rate:.5f}")
    assert amp.max() <= 1e-4 and width.max() <= 1e-4 and center.max() <= 1e-4
    assert chi.max() <= 1e-5


@pytest.mark.parametrize("mapping", MAPPINGS)
@pytest.mark.parametrize("numerics", NUMERICS)
def test_failure_isolation(hip, golden, mapping, numerics):
    """tests/metal_lm_checks.hpp:59-128 and cooperative_checks.hpp:203-257;
    also a 16-lane request for M=1024, which no register tile holds, so the
    points stream."""
    g = golden["failures"]
    x, y, p0 = datasets.isolation_fixture()
    kw = dict(mapping=mapping, numerics=numerics)
    r = hip.fit(x, y, "lorentzian", p0, None, "hip", **kw)
    assert list(r.status) == [2, 0, 3] and list(r.iterations) == [1, 0, 0]
    assert r.chi2[0] == 3 and np.isnan(r.chi2[2])
    np.testing.assert_array_equal(r.params, g["iso_0_params"])
    r = hip.fit(x, y, "lorentzian", p0, None, "hip", max_iterations=0, **kw)
    assert r.status[0] == 1 and r.chi2[0] == 3
    r = hip.fit(x, y, "lorentzian", p0, ([-np.inf, 0, 1], [np.inf, 0, 1]), "hip", **kw)
    assert r.status[0] == 0 and r.params[0, 1] == 0 and r.params[0, 2] == 1
    r = hip.fit(x, y, "lorentzian", p0, ([-np.inf, 0, 1], [1.5, 0, 1]), "hip", **kw)
    assert np.all(r.status == 0) and np.all(r.params[:, 0] == np.float32(1.5))
    bad_x = x.copy()
    bad_x[0] = np.nan
    assert np.all(hip.fit(bad_x, y, "lorentzian", p0, None, "hip", **kw).status == 3)
    bad_y = y.copy()
    bad_y[1, 1] = np.nan
    r = hip.fit(x, bad_y, "lorentzian", p0, None, "hip", **kw)
    assert r.status[0] == 2 and r.status[1] == 3
    empty = hip.fit(np.linspace(0, 1, 32), np.empty((0, 32)), "gaussian", None, None, "hip", **kw)
    assert empty.params.shape == (0, 3) and empty.status.shape == (0,)
    if mapping == "group16":
        # Beyond the largest register tile the 16 lanes read the points from
        # memory: the same result as asking for that form.
        wide = np.linspace(-1, 1, 1024)
        y_wide = np.ones((2, 1024))
        a = hip.fit(wide, y_wide, "gaussian", [1, 0, 1], None, "hip", **kw)
        kw_streamed = dict(kw, mapping="group16_streamed")
        b = hip.fit(wide, y_wide, "gaussian", [1, 0, 1], None, "hip", **kw_streamed)
        fields_bitwise_equal(a, b, "group16 M=1024")


@pytest.mark.parametrize("model", list(P))
def test_pathology_matrix(hip, golden, model):
    """tests/numerics_checks.hpp:280-350: degenerate, non-finite, edge,
    sub-spacing, ordinary and overflow rows stay isolated in every form."""
    g = golden["failures"]
    x, inits, _ = datasets.pathology_fixture(model)
    y = g[f"path_{model}_y"]
    for mapping in MAPPINGS:
        for numerics in (0, 1, 2, 3, 5, 7):
            if mapping in ("group4_streamed", "group8", "group16", "group32", "group256") and \
                    numerics not in (0, 5):
                continue
            r = hip.fit(x, y, model, inits, None, "hip", mapping=mapping, numerics=numerics)
            assert list(r.status) == [4, 4, 3, 3, 3, 0, 5, 0, 3], (mapping, numerics, r.status)
            assert np.all(np.isnan(r.chi2[2:5]))
            for i in (0, 1, 5, 6, 7):
                assert np.isfinite(r.chi2[i]) and np.all(np.isfinite(r.params[i]))
    bad_x = x.copy()
    bad_x[17] = np.nan
    for mapping in MAPPINGS:
        assert np.all(hip.fit(bad_x, y, model, inits, None, "hip", mapping=mapping).status == 3)


def test_row_isolation_api(hip):
    """python/tests/test_api.py:69-76."""
    x, y, p0 = datasets.api_isolation_fixture(np.float32)
    r = hip.fit(x, y, "lorentzian", p0[0], None, "hip")
    np.testing.assert_array_equal(r.status, [0, 3, 3, 4, 0])
    c = hip.fit(x, y[[0]], "lorentzian", p0[0], ([0, .2, .7], [np.inf, .2, .7]), "hip")
    np.testing.assert_allclose(c.params[0], [2, .2, .7], atol=2e-4)


@pytest.mark.parametrize("model", list(P))
def test_uncertainties_match_oracle(hip, oracle_port, model):
    """tests/uncertainty_checks.hpp:104-190: covariance at the accepted point
    within 1e-9 + 2e-4·|cpu| of the FP32 oracle applied to the GPU's own fit."""
    x, y, p0, bounds, _ = datasets.comparison_set(model, 64, 129)
    x, y = x.astype(np.float32), y.astype(np.float32)
    for mapping in MAPPINGS:
        basic = hip.fit(x, y, model, p0, bounds, "hip", mapping=mapping)
        full = hip.fit(x, y, model, p0, bounds, "hip", mapping=mapping, uncertainties=True)
        for a, b in zip(basic, full[:4]):
            assert np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                                  np.ascontiguousarray(b).view(np.uint8)), "fit bytes changed"
        for i in range(64):
            cov, se, us = oracle_port.uncertainty(model, x, full.params[i], int(full.status[i]),
                                                  float(full.chi2[i]), bounds, dtype=np.float32)
            assert full.uncertainty_status[i] == us, (mapping, i)
            if us == 0:
                np.testing.assert_allclose(full.covariance[i], cov, rtol=2e-4, atol=1e-9)
                np.testing.assert_allclose(full.stderr[i], se, rtol=2e-4, atol=1e-8)
        np.testing.assert_array_equal(full.covariance, full.covariance.transpose(0, 2, 1))


def test_uncertainty_statuses(hip):
    """python/tests/test_uncertainty.py:67-98."""
    x = np.linspace(-5, 5, 129)
    y = datasets.lorentzian(x, 2, .2, .7)
    initial = [2, .2, .7]
    failed = y.copy()
    failed[17] = np.nan
    r = hip.fit(x, failed, p0=initial, device="hip", uncertainties=True)
    assert r.status[0] == 3 and r.uncertainty_status[0] == 1
    assert np.all(np.isnan(r.covariance)) and np.all(np.isnan(r.stderr))
    small_x = np.array([-1., 0., 1.])
    small = hip.fit(small_x, datasets.lorentzian(small_x, 2, 0, 2), p0=[2, 0, 2], device="hip",
                    uncertainties=True)
    assert small.uncertainty_status[0] == 2
    singular = hip.fit(np.zeros(20), np.full(20, 2.), p0=[2, 0, 1], device="hip",
                       uncertainties=True)
    assert singular.status[0] == 0 and singular.uncertainty_status[0] == 3
    boundary = hip.fit(x, y, p0=initial, bounds=([2, -5, .01], [np.inf, 5, 5]), device="hip",
                       uncertainties=True)
    assert boundary.uncertainty_status[0] == 4
    noisy = y + np.random.default_rng(101).normal(0, .03, len(x))
    fixed = hip.fit(x, noisy, p0=initial, bounds=([0, .2, .7], [np.inf, .2, .7]), device="hip",
                    uncertainties=True)
    assert fixed.uncertainty_status[0] == 0
    np.testing.assert_array_equal(fixed.covariance[0, 1:], 0)
    np.testing.assert_array_equal(fixed.stderr[0, 1:], 0)
    shape = datasets.lorentzian(x, 1, .2, .7)
    expected = fixed.chi2[0] / (len(x) - 1) / np.sum(shape ** 2)
    np.testing.assert_allclose(fixed.covariance[0, 0, 0], expected, rtol=1e-4, atol=1e-12)
    empty = hip.fit(x, np.empty((0, len(x))), device="hip", uncertainties=True)
    assert empty.covariance.shape == (0, 3, 3) and empty.stderr.shape == (0, 3)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 255, 256, 257])
def test_batch_tails(hip, oracle_port, n):
    x, y, p0, bounds, _ = datasets.comparison_set("lorentzian", 257, 129)
    x, y = x.astype(np.float32), y[:n].astype(np.float32)
    ref = oracle_port.fit(x, y, "lorentzian", p0, bounds)
    for mapping in MAPPINGS:
        assert_fp32_gate(hip.fit(x, y, "lorentzian", p0, bounds, "hip", mapping=mapping), ref,
                         f"N={n} {mapping}")


@pytest.mark.parametrize("m", [5, 31, 33, 64, 65, 2047, 2049, 4096])
def test_point_count_edges(hip, oracle_port, m):
    """Register tiles of every size, the partial last tile and the streamed form."""
    x, y, p0, bounds, _ = datasets.comparison_set("pseudo_voigt", 48, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    ref = oracle_port.fit(x, y, "pseudo_voigt", p0, bounds, threads=8)
    for mapping in mappings_for(m):
        assert_fp32_gate(hip.fit(x, y, "pseudo_voigt", p0, bounds, "hip", mapping=mapping), ref,
                         f"M={m} {mapping}")


@pytest.mark.parametrize("model", list(P))
@pytest.mark.parametrize("g,m", [(128, 129), (128, 1500), (256, 1024), (256, 2048), (256, 4096),
                                 (256, 9000)])
def test_workgroup_per_fit(hip, oracle_port, model, g, m):
    """A workgroup of 2 or 4 waves per fit (the reference's threadgroup-per-fit
    form, src/metal/cooperative.metal:16-32, 338-349): partials reduced within
    each wave, then across the waves through LDS. Within the FP32 gate of the
    oracle in both numerics, the register tile (8 or 32 points per lane) gives
    the streamed form's bytes, and uncertainties leave the fit's bytes alone.
    (Its sums run in another order than a 64-lane group's, so its bytes are
    not a 64-lane group's.)"""
    x, y, p0, bounds, _ = datasets.comparison_set(model, 33, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    ref = oracle_port.fit(x, y, model, p0, bounds, threads=8)
    for numerics in NUMERICS:
        a = hip.fit(x, y, model, p0, bounds, "hip", mapping=f"group{g}", numerics=numerics)
        assert_fp32_gate(a, ref, f"{model} M={m} group{g} numerics={numerics}")
        b = hip.fit(x, y, model, p0, bounds, "hip", mapping=f"group{g}_streamed",
                    numerics=numerics)
        fields_bitwise_equal(a, b, f"{model} M={m} group{g} tile vs streamed")
    full = hip.fit(x, y, model, p0, bounds, "hip", mapping=f"group{g}", uncertainties=True)
    fields_bitwise_equal(hip.FitResult(*full[:4]),
                         hip.fit(x, y, model, p0, bounds, "hip", mapping=f"group{g}"),
                         f"{model} M={m} group{g} uncertainties")
    np.testing.assert_array_equal(full.covariance, full.covariance.transpose(0, 2, 1))
    guessed = hip.fit(x, y, model, None, bounds, "hip", mapping=f"group{g}")
    assert_fp32_gate(guessed, oracle_port.fit(x, y, model, None, bounds, threads=8),
                     f"{model} M={m} group{g} from the device guess")


@pytest.mark.parametrize("model,m,n", [("pseudo_voigt", 4096, 64), ("two_lorentzian", 2048, 48),
                                       ("gaussian", 300, 2000)])
def test_latency_mapping_is_the_mapping_it_names(hip, oracle_port, model, m, n):
    """mapping="latency" runs the mapping plan(..., fits=N) names — the same
    bytes as asking for it — within the FP32 gate."""
    x, y, p0, bounds, _ = datasets.comparison_set(model, n, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    lanes, tile = hip.plan(model, m, mapping="latency", fits=n)
    assert lanes >= 64
    name = f"group{lanes}" + ("" if tile else "_streamed")
    a = hip.fit(x, y, model, None, bounds, "hip", mapping="latency")
    b = hip.fit(x, y, model, None, bounds, "hip", mapping=name)
    fields_bitwise_equal(a, b, f"{model} M={m} N={n} latency = {name}")
    assert_fp32_gate(a, oracle_port.fit(x, y, model, None, bounds, threads=8),
                     f"{model} M={m} N={n} latency")


@pytest.mark.parametrize("unc,given", [(False, False), (True, True)])
def test_host_arrays_stream_in_chunks(hip, unc, given):
    """bf_hip_fit (host arrays) takes the batch through the device in chunks of
    about 64 MiB of y on two streams (copy-in of one chunk beside the fit of
    another): the same bytes as one device call on the whole batch."""
    import ctypes
    from batchfit_amd import _native
    from batchfit_amd.ragged import DeviceArray
    x, y, bounds, _ = datasets.config2(rows=40000)  # 3 chunks of 16,384 fits at M = 1024
    n, m = y.shape
    p0 = np.tile(np.array([2.0, 0.0, 1.0, 0.5], np.float32), (n, 1)) if given else None
    host = hip.fit(x, y, "pseudo_voigt", p0, bounds, "hip", uncertainties=unc)
    lib = _native.library()
    cfg = _native.default_config()
    cfg.fits, cfg.points, cfg.numerics = n, m, hip.NUMERICS_DEFAULT
    for k in range(4):
        cfg.lower[k], cfg.upper[k] = bounds[0][k], bounds[1][k]
    dx, dy = DeviceArray.of(x), DeviceArray.of(y)
    dp0 = DeviceArray.of(p0) if given else None
    outs = [DeviceArray(n * 16), DeviceArray(n * 4), DeviceArray(n * 4), DeviceArray(n * 4)]
    if unc:
        outs += [DeviceArray(n * 64), DeviceArray(n * 16), DeviceArray(n * 4)]
    ptrs = [o.ptr for o in outs] + ([] if unc else [None] * 3)
    err = _native.error_buffer()
    _native.check(lib.bf_hip_fit_device(2, ctypes.byref(cfg), None, dx.ptr, dy.ptr,
                                        dp0.ptr if given else None, *ptrs, err, len(err)), err)
    shapes = [(n, 4), (n,), (n,), (n,), (n, 4, 4), (n, 4), (n,)]
    types = [np.float32, np.int32, np.uint32, np.float32, np.float32, np.float32, np.int32]
    dev = [o.download(np.empty(sh, t)) for o, sh, t in zip(outs, shapes, types)]
    order = [host.params, host.status, host.iterations, host.chi2] + (
        [host.covariance, host.stderr, host.uncertainty_status] if unc else [])
    for a, b in zip(order, dev):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint8), b.view(np.uint8))
    for o in outs + [dx, dy] + ([dp0] if given else []):
        o.free()


@pytest.mark.parametrize("model", ["gaussian", "lorentzian", "pseudo_voigt"])
@pytest.mark.parametrize("m", [10250, 16384, 65536])
def test_guess_long_spectra(hip, oracle_port, model, m):
    """The single-peak guess has no limit on M: edge windows over 1024 values
    are searched with their keys read from memory, and the row streams. Bit
    for bit the oracle's guess (initial_guess.hpp keeps its scratch on the
    heap, :143); fits from it within the FP32 gate."""
    x, y, p0, bounds, _ = datasets.comparison_set(model, 8, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    got = device_guess(x, y, model)
    ref = oracle_port.guess(x, y, model, threads=8)
    for a, b in zip(got, ref):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                              np.ascontiguousarray(b).view(np.uint8)), (model, m)
    assert np.all(ref[2] == 0)
    q = (np.round(y * 50) / 50).astype(np.float32)
    got = device_guess(x, q, model)
    ref = oracle_port.guess(x, q, model, threads=8)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b, equal_nan=True), (model, m, "quantised")
    fit = hip.fit(x, y, model, None, bounds, "hip")
    assert_fp32_gate(fit, oracle_port.fit(x, y, model, None, bounds, threads=8),
                     f"{model} M={m} guessed")


def test_guess_form_errors(hip):
    """The one-wave form holds at most 15,886 two-peak points in LDS (16,384
    single-peak); unknown forms are argument errors."""
    from batchfit_amd import _native
    x = np.linspace(-5, 5, 15887).astype(np.float32)
    y = np.ones((2, x.size), np.float32)
    with pytest.raises(ValueError, match="holds M up to 15886 points"):
        device_guess(x, y, "two_lorentzian", GUESS_WAVE)
    x1 = np.linspace(-5, 5, 16385).astype(np.float32)
    with pytest.raises(ValueError, match="holds M up to 16384 points"):
        device_guess(x1, np.ones((2, x1.size), np.float32), "lorentzian", GUESS_WAVE)
    with pytest.raises(ValueError, match="unknown guess form"):
        device_guess(x, y, "two_lorentzian", 3)


def peak_scan_rows(m, rng):
    """Rows aimed at the single-peak scan's corners: the largest smoothed value
    at j = 1 or M − 2 with the end point above it (the reference then takes
    another local maximum) or below it, flat tops (ties: the first index
    wins), a level of −0 (the signal reads the grid), non-finite observations
    and a signal that overflows, no positive maximum, and plain noisy peaks."""
    x = np.linspace(-5, 5, m)
    bump = np.exp(-0.5 * (x / 0.4) ** 2)
    rows = []
    ramp = np.linspace(10, 0, m)
    rows += [ramp + 0.5 * bump, ramp[::-1] + 0.5 * bump, ramp + 3 * bump, ramp[::-1] + 3 * bump]
    r = np.linspace(4, 0, m)
    r[:3] = [5, 10, 4]
    rows += [r, r[::-1].copy()]
    flat = np.zeros(m)
    flat[m // 3: m // 3 + 5] = 2.0
    rows += [flat, np.round(3 * bump, 1)]
    neg = np.full(m, -0.0)
    neg[m // 2 - 2: m // 2 + 3] = [0.5, 1.5, 2.0, 1.5, 0.5][: min(5, m)]
    rows += [neg, -bump]
    big = np.full(m, -3e38)
    big[m // 2] = 3e38
    nan = 1 + bump
    nan[m // 2] = np.nan
    inf = 1 + bump
    inf[0] = np.inf
    rows += [big, nan, inf]
    rows += [1 + bump + 0.05 * rng.standard_normal(m) for _ in range(5)]
    return x.astype(np.float32), np.array(rows, np.float32)


@pytest.mark.parametrize("m", [5, 6, 7, 8, 63, 64, 65, 67, 68, 255, 256, 257, 1023, 1024, 1028])
def test_guess_peak_scan_corners(hip, oracle_port, m):
    rng = np.random.default_rng(m)
    x, y = peak_scan_rows(m, rng)
    for model in ("gaussian", "lorentzian", "pseudo_voigt"):
        got = device_guess(x, y, model)
        ref = oracle_port.guess(x, y, model)
        for a, b in zip(got, ref):
            assert np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                                  np.ascontiguousarray(b).view(np.uint8)), (model, m, a, b)
