"""GPU tests of ragged batches (bf_hip_ragged_*): fits of mixed model, M and
bounds in one captured launch.

* Same alone and in the mix: every fit's bytes equal those of the same fit in
  a plan of its own batch (the mapping depends on (model, M) only).
* Parity: each batch within the reference's FP32 GPU gate of the FP32 oracle
  (params ≤ 2e-4 + 2e-4·|cpu|, χ² ≤ 1e-6 + 2e-5·|cpu|, equal status).
* Graph replays: the captured graph replayed several times, with eager work
  between replays, gives the eager result every time.
"""
import numpy as np
import pytest

import datasets

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                          np.ascontiguousarray(b).view(np.uint8))


def mixed_batches():
    out = []
    for model in datasets.MODELS:
        for m, n in ((64, 37), (129, 40), (300, 21), (1024, 19), (1500, 9), (2500, 5)):
            x, y, p0, bounds, _ = datasets.comparison_set(model, n, m)
            out.append((x.astype(np.float32), y.astype(np.float32), model,
                        None if m in (129, 1024) else p0, bounds))
    return out


def test_ragged_matches_oracle_and_is_batch_independent(hip, oracle_port):
    from batchfit_amd.ragged import fit_ragged
    batches = mixed_batches()
    together = fit_ragged(batches)
    for (x, y, model, p0, bounds), got in zip(batches, together):
        ref = oracle_port.fit(x, y, model, p0, bounds, threads=8)
        label = f"{model} M={x.size}"
        assert np.array_equal(got.status, ref.status), label
        ok = ref.status == 0
        assert np.all(np.abs(got.params[ok] - ref.params[ok])
                      <= 2e-4 + 2e-4 * np.abs(ref.params[ok])), label
        assert np.all(np.abs(got.chi2[ok] - ref.chi2[ok]) <= 1e-6 + 2e-5 * np.abs(ref.chi2[ok]))
        alone = fit_ragged([(x, y, model, p0, bounds)])[0]
        for a, b in zip(alone, got):
            assert same(a, b), f"{label}: result depends on the rest of the batch"
        one = fit_ragged([(x, y[:1], model, p0, bounds)])[0]
        for a, b in zip(one, got):
            assert same(a, b[:1]), label


def test_ragged_equals_uniform_path(hip):
    """The uniform entry uses the same tuned (model, M) table: same bytes."""
    from batchfit_amd.ragged import fit_ragged
    for model in ("lorentzian", "pseudo_voigt"):
        for m in (100, 700, 1800):
            x, y, p0, bounds, _ = datasets.comparison_set(model, 33, m)
            x, y = x.astype(np.float32), y.astype(np.float32)
            a = hip.fit(x, y, model, p0, bounds, "hip")
            b = fit_ragged([(x, y, model, p0, bounds)])[0]
            for u, v in zip(a, b):
                assert same(u, v), (model, m)


@pytest.mark.parametrize("shift", [0.0, 2.0e6])
def test_fixed_baseline_doublets_equal_full_form(hip, oracle_port, shift):
    """Doublets with the baseline fixed by equal bounds: the pool kernels that
    sum only the six free parameters' rows run where every grid of the batch
    lies within |x| <= 1e6, the full kernels elsewhere (a uniform batch on a
    wider grid, a ragged batch holding one). The same bytes every way."""
    from batchfit_amd.ragged import fit_ragged
    x, y, _, _, _ = datasets.comparison_set("two_lorentzian", 300, 900)
    x = (shift + 80.0 * x).astype(np.float32)  # spacing 0.89, above an ulp at 2e6
    y = y.astype(np.float32)
    assert np.all(np.diff(x) > 0)
    lower = np.array([0, shift - 400, .8, 0, shift - 400, .8, 0.3, 0], np.float64)
    upper = np.array([np.inf, shift + 400, 400, np.inf, shift + 400, 400, 0.3, 0], np.float64)
    bounds = (lower, upper)
    a = hip.fit(x, y, "two_lorentzian", None, bounds, "hip")
    b = fit_ragged([(x, y, "two_lorentzian", None, bounds)])[0]
    for u, v in zip(a, b):
        assert same(u, v), shift
    # A ragged batch that also holds a grid beyond 1e6 runs its fixed-baseline
    # classes in the full form (the LM and the uncertainty kernels).
    xl, yl, p0l, bl, _ = datasets.comparison_set("lorentzian", 5, 400)
    xl = (2.0e6 + 80.0 * xl).astype(np.float32)
    bl = (np.array([0, 2.0e6 - 400, .8]), np.array([np.inf, 2.0e6 + 400, 400]))
    c = fit_ragged([(x, y, "two_lorentzian", None, bounds),
                    (xl, yl.astype(np.float32), "lorentzian", None, bl)])[0]
    for u, v in zip(a, c):
        assert same(u, v), (shift, "with a wide grid in the batch")
    au = hip.fit(x, y, "two_lorentzian", None, bounds, "hip", uncertainties=True)
    bu = fit_ragged([(x, y, "two_lorentzian", None, bounds)], uncertainties=True)[0]
    cu = fit_ragged([(x, y, "two_lorentzian", None, bounds),
                     (xl, yl.astype(np.float32), "lorentzian", None, bl)], uncertainties=True)[0]
    for u, v, w in zip(au, bu, cu):
        assert same(u, v) and same(u, w), (shift, "uncertainties")
    assert np.sum(au.uncertainty_status == 0) > 100, shift
    # The grid is 80x the unit-scale sets', so positions and widths are too:
    # statuses and chi2 against the oracle (the unit-scale parameter gate of
    # the other tests holds this path at M = 2048 on the doublet workload).
    ref = oracle_port.fit(x, y, "two_lorentzian", None, bounds, threads=8)
    ok = (ref.status == 0) & (a.status == 0)
    assert np.mean(a.status == ref.status) >= 0.99 and ok.sum() > 100
    assert np.all(np.abs(a.chi2[ok] - ref.chi2[ok]) <= 1e-4 * np.abs(ref.chi2[ok]))


def test_ragged_long_spectra_from_guess(hip):
    """Guessed fits beyond the one-wave guess's LDS (M > 16,384 points) take
    the wide guess in a ragged batch: the same bytes as the uniform entry
    (whose single-peak guess is the streamed group form), beside short fits."""
    from batchfit_amd.ragged import fit_ragged
    batches = []
    for model, m, n in (("two_lorentzian", 16384, 5), ("lorentzian", 20000, 4),
                        ("gaussian", 17000, 3), ("pseudo_voigt", 300, 30)):
        x, y, p0, bounds, _ = datasets.comparison_set(model, n, m)
        batches.append((x.astype(np.float32), y.astype(np.float32), model, None, bounds))
    together = fit_ragged(batches)
    for (x, y, model, p0, bounds), got in zip(batches, together):
        ref = hip.fit(x, y, model, None, bounds, "hip")
        for a, b in zip(got, ref):
            assert same(a, b), (model, x.size)
        assert np.all(ref.status == 0), (model, x.size)


@pytest.mark.parametrize("model,m", [("pseudo_voigt", 900), ("two_lorentzian", 700),
                                     ("gaussian", 3000)])
def test_ragged_equals_uniform_large_streamed_class(hip, model, m):
    """A streamed class large enough that every wave refills its pool many
    times, from the device guess: the same bytes as the uniform launch."""
    from batchfit_amd.ragged import fit_ragged
    x, y, p0, bounds, _ = datasets.comparison_set(model, 12000, m)
    x, y = x.astype(np.float32), y.astype(np.float32)
    a = hip.fit(x, y, model, None, bounds, "hip")
    b = fit_ragged([(x, y, model, None, bounds)])[0]
    for u, v in zip(a, b):
        assert same(u, v), (model, m)
    # With uncertainties (a second launch after the retiring one, both paths).
    a = hip.fit(x[:], y[:3000], model, None, bounds, "hip", uncertainties=True)
    b = fit_ragged([(x, y[:3000], model, None, bounds)], uncertainties=True)[0]
    for name, u, v in zip(a._fields, a, b):
        assert same(u, v), (model, m, name)


def test_ragged_split_runs_equal_uniform(hip):
    """A class heavier than an eighth of the batch and than 2^26 weighted
    points (40,000 two-peak fits of M = 2048: 2.9e8) is cut into several runs
    over consecutive slices of its order; every fit keeps the bytes of the
    uniform launch, and the small batch beside it too."""
    from batchfit_amd.ragged import fit_ragged
    x, y, _, bounds, _ = datasets.comparison_set("two_lorentzian", 40000, 2048)
    x, y = x.astype(np.float32), y.astype(np.float32)
    xs, ys, _, bs, _ = datasets.comparison_set("gaussian", 50, 129)
    xs, ys = xs.astype(np.float32), ys.astype(np.float32)
    big, small = fit_ragged([(x, y, "two_lorentzian", None, bounds),
                             (xs, ys, "gaussian", None, bs)])
    for u, v in zip(hip.fit(x, y, "two_lorentzian", None, bounds, "hip"), big):
        assert same(u, v)
    for u, v in zip(hip.fit(xs, ys, "gaussian", None, bs, "hip"), small):
        assert same(u, v)


def test_ragged_uncertainties_and_failures(hip, oracle_port):
    from batchfit_amd.ragged import fit_ragged
    x, y, p0, bounds, _ = datasets.comparison_set("gaussian", 20, 129)
    x, y = x.astype(np.float32), y.astype(np.float32)
    y[3, 7] = np.nan
    y[4] = 1.0
    xi, yi, pi = datasets.isolation_fixture()
    res = fit_ragged([(x, y, "gaussian", p0, bounds), (xi, yi, "lorentzian", pi, None)],
                     uncertainties=True)
    g = res[0]
    assert g.status[3] == 3 and g.status[4] == 4
    assert g.covariance.shape == (20, 3, 3)
    for i in range(20):
        cov, se, us = oracle_port.uncertainty("gaussian", x, g.params[i], int(g.status[i]),
                                              float(g.chi2[i]), bounds, dtype=np.float32)
        assert g.uncertainty_status[i] == us
        if us == 0:
            np.testing.assert_allclose(g.covariance[i], cov, rtol=2e-4, atol=1e-9)
    iso = res[1]
    assert list(iso.status) == [2, 0, 3] and list(iso.iterations) == [1, 0, 0]


def test_graph_replays_match_eager(hip):
    from batchfit_amd import ragged
    # Every fit starts from the device guess, as in the plan built below.
    batches = [(x, y, model, None, bounds) for x, y, model, _, bounds in mixed_batches()[:8]]
    first = ragged.fit_ragged(batches)
    # Build one plan and replay its graph several times with eager work between.
    import batchfit_amd._native as native
    xs, ys, descs, grids, configs = [], [], [], [], []
    x_off = y_off = row = 0
    for b, (x, y, model, p0, bounds) in enumerate(batches):
        n, m = y.shape
        cfg = native.default_config()
        cfg.numerics = 5
        for k in range(len(bounds[0])):
            cfg.lower[k], cfg.upper[k] = float(bounds[0][k]), float(bounds[1][k])
        configs.append(cfg)
        grids.append((x_off, m, 0))
        d = np.zeros(n, ragged.FIT_DESC)
        d["y_offset"] = y_off + np.arange(n) * m
        d["points"], d["grid"], d["model"], d["config"] = m, b, native.MODELS[model], b
        d["p0_row"], d["out_row"] = ragged.NO_P0, row + np.arange(n)
        descs.append(d)
        xs.append(x)
        ys.append(y.reshape(-1))
        x_off, y_off, row = x_off + m, y_off + n * m, row + n
    plan = ragged.RaggedPlan(np.concatenate(descs), np.array(grids, ragged.GRID_DESC), configs)
    dx, dy = ragged.DeviceArray.of(np.concatenate(xs)), ragged.DeviceArray.of(np.concatenate(ys))
    out = ragged.allocate_outputs(row, False)
    expected = np.concatenate([r.params for r in first], axis=None)
    for replay in range(4):
        plan.run(dx, dy, None, out)
        host = ragged.download_outputs(out, row)
        got = np.concatenate([host["params"][s:s + n, :p].reshape(-1) for s, n, p in
                              _shapes(batches)])
        assert same(got, expected), f"replay {replay}"
        eager = hip.fit(batches[0][0], batches[0][1], batches[0][2], None, batches[0][4], "hip")
        assert eager.params.shape[0] == batches[0][1].shape[0]
    plan.close()


def _shapes(batches):
    P = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8}
    row = 0
    for x, y, model, p0, bounds in batches:
        yield row, y.shape[0], P[model]
        row += y.shape[0]
