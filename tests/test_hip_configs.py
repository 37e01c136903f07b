"""GPU fits of the README quickstart (Lorentzian, N=1000, M=129) and of the
overlapping-doublet workload (two_lorentzian, M=2048, 5 % degenerate rows,
uncertainties) against the CPU oracle on identical inputs.

Gates, written here:
  * FP32 gate (tests/metal_lm_checks.hpp:41-45 of the reference): status
    equal; converged params |gpu−cpu| ≤ 2e-4 + 2e-4·|cpu|, χ² ≤ 1e-6 +
    2e-5·|cpu|.
  * FP64 gate: status agreement ≥ 99.9 %; converged rows amplitude and width
    relative error ≤ 1e-4, centre error ≤ 1e-4·width, χ² relative ≤ 1e-5.
  * Doublets: a doublet whose peaks overlap is ill-conditioned — the oracle's
    own FP32 results move when only its summation order changes (plain sums
    vs Kahan sums agree on 99.6 % of statuses, and 97 % of converged rows
    stay within the FP32 gate). The GPU sums in another order again, so the
    FP32 gate is applied as: the GPU's agreement with the FP32 oracle is at
    least the oracle's agreement with itself under the other summation, less
    0.5 percentage points of statuses and 1 point of rows within the gate.
    Against FP64 the reference's own FP32 path agrees on about 82 % of
    statuses only (the FP64 defaults stop at 1e-10/1e-10/1e-12 where FP32
    stops at 1e-5/1e-6/1e-6, and where a width collapses below the spacing
    FP32 stops at the step test with unresolved_peak while FP64 runs to
    max_iterations), so the 99.9 % FP64 gate cannot hold on this workload;
    the test requires the GPU to agree with FP64 at least as well as the FP32
    oracle does (less 0.5 points), prints every disagreement with the values
    of the stop tests against their thresholds, and prints the agreement with
    FP64 run at the FP32 tolerances.
Structure is exact: every constant or NaN row gets invalid_input (the guess
fails, as in the reference with p0=None), the fixed baseline's covariance
rows and columns and standard errors are exactly zero, uncertainty statuses
equal the FP64 oracle's evaluated at the GPU's own fit.
"""
from collections import Counter

import numpy as np
import pytest

import datasets

pytestmark = pytest.mark.gpu

STATUS = ["converged", "max_iterations", "singular_matrix", "invalid_input", "degenerate_data",
          "unresolved_peak"]
FIRED = {0: "none", 1: "gradient", 2: "step", 3: "cost"}


def fp32_gate_rows(got_p, got_c, ref_p, ref_c):
    """Rows within the FP32 parameter and χ² gate."""
    dp = np.abs(got_p - ref_p) <= 2e-4 + 2e-4 * np.abs(ref_p)
    dc = np.abs(got_c - ref_c) <= 1e-6 + 2e-5 * np.abs(ref_c)
    return dp.all(axis=1) & dc


def fp64_errors(got, ref, rows):
    g, r = got.params[rows].astype(np.float64), ref.params[rows]
    amp = np.abs(g[:, 0] - r[:, 0]) / np.abs(r[:, 0])
    width = np.abs(g[:, 2] - r[:, 2]) / np.abs(r[:, 2])
    centre = np.abs(g[:, 1] - r[:, 1]) / np.abs(r[:, 2])
    chi = np.abs(got.chi2[rows] - ref.chi2[rows]) / ref.chi2[rows]
    return amp, width, centre, chi


def describe(stops):
    """One oracle fit's last stop tests against their thresholds."""
    g, gt, s, st, red, ct, pr, pt, w, lam, fired = stops
    return (f"ended by {FIRED[int(fired)]}: |g| {g:.3g}/tol {gt:.3g}, step {s:.3g}/thr {st:.3g}, "
            f"|dchi2| {red:.3g} pred {pr:.3g}/thr {ct:.3g},{pt:.3g}, width*f/spacing {w:.4g}, "
            f"lambda {lam:.3g}")


def width_ratio(params, spacing):
    return np.minimum(np.abs(params[:, 2]), np.abs(params[:, 5])) * 2 / spacing


# ---------------------------------------------------------------------------
# Config 1: the README quickstart, float32 and float64 inputs.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("numerics", [0, 5])
def test_config1_float32_matches_reference_outputs(hip, golden, numerics):
    """Float32 inputs against the reference's own float32 outputs for the same
    arrays (tests/golden/golden_config1.npz, written by oracle/_ref), FP32 gate."""
    g = golden["config1"]
    x, y, bounds = datasets.config1(np.float32)
    got = hip.fit(x, y, "lorentzian", None, bounds, "hip", numerics=numerics)
    assert np.array_equal(got.status, g["c1_f32_status"])
    ok = g["c1_f32_status"] == 0
    assert ok.all()
    within = fp32_gate_rows(got.params, got.chi2, g["c1_f32_params"], g["c1_f32_chi2"])
    assert within.all(), np.nonzero(~within)[0][:10]
    assert np.all(got.iterations <= 100)


def test_config1_float64_inputs_fp64_gate(hip, golden, oracle_port):
    """Float64 inputs through device="hip" (converted to float32, float32
    results, as the reference's explicit GPU device) against the reference's
    own float64 outputs on the float64 arrays, and against the FP64 oracle on
    the converted float32 arrays; FP64 gate in both cases."""
    g = golden["config1"]
    x, y, bounds = datasets.config1(np.float64)
    got = hip.fit(x, y, "lorentzian", None, bounds, "hip")
    assert got.params.dtype == np.float32 and got.chi2.dtype == np.float32
    same32 = hip.fit(x.astype(np.float32), y.astype(np.float32), "lorentzian", None, bounds, "hip")
    for a, b in zip(got, same32):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    ref64_same = oracle_port.fit(x.astype(np.float32).astype(np.float64),
                                 y.astype(np.float32).astype(np.float64), "lorentzian", None,
                                 bounds, threads=8)
    for label, status, params, chi2 in (
            ("reference float64 outputs", g["c1_f64_status"], g["c1_f64_params"], g["c1_f64_chi2"]),
            ("FP64 oracle on the float32 arrays", ref64_same.status, ref64_same.params,
             ref64_same.chi2)):
        ref = type(got)(params, status, None, chi2)
        agree = np.mean(got.status == status)
        both = (got.status == 0) & (status == 0)
        amp, width, centre, chi = fp64_errors(got, ref, both)
        print(f"config 1 vs {label}: status agreement {agree:.4f}; max relative error "
              f"amplitude {amp.max():.3g}, width {width.max():.3g}, centre/width "
              f"{centre.max():.3g}, chi2 {chi.max():.3g}")
        assert agree >= 0.999
        assert amp.max() <= 1e-4 and width.max() <= 1e-4 and centre.max() <= 1e-4
        assert chi.max() <= 1e-5


# ---------------------------------------------------------------------------
# Config 3: overlapping doublets with uncertainties.
# ---------------------------------------------------------------------------
ROWS = 8192


@pytest.fixture(scope="module")
def doublets(oracle_port):
    from oracle import oracle
    x, y, bounds, truth, constant, with_nan = datasets.config3(rows=ROWS)
    kw = dict(threads=16)
    x64, y64 = x.astype(np.float64), y.astype(np.float64)
    return {
        "x": x, "y": y, "bounds": bounds, "constant": constant, "with_nan": with_nan,
        "f32": oracle_port.fit(x, y, "two_lorentzian", None, bounds, **kw),
        "f32k": oracle_port.fit(x, y, "two_lorentzian", None, bounds,
                                options=oracle.make_options(compensated=True), **kw),
        "f64": oracle_port.fit(x64, y64, "two_lorentzian", None, bounds, **kw),
        "f64_at_f32_tol": oracle_port.fit(
            x64, y64, "two_lorentzian", None, bounds,
            options=oracle.make_options(gradient_tolerance=1e-5, step_tolerance=1e-6,
                                        cost_tolerance=1e-6), **kw),
        "stops32": oracle_port.fit_stops(x, y, "two_lorentzian", None, bounds, **kw),
        "stops64": oracle_port.fit_stops(x64, y64, "two_lorentzian", None, bounds, **kw),
    }


@pytest.mark.parametrize("numerics", [0, 5])
def test_config3_doublets_fp32_gate(hip, doublets, numerics):
    d = doublets
    x, y, bounds = d["x"], d["y"], d["bounds"]
    got = hip.fit(x, y, "two_lorentzian", None, bounds, "hip", numerics=numerics,
                  uncertainties=True)
    ref, other = d["f32"], d["f32k"]
    # Degenerate rows: the guess fails (no peak above the noise / non-finite
    # data), so the fit starts from NaN parameters: invalid_input, NaN χ².
    degenerate = d["constant"] | d["with_nan"]
    assert degenerate.sum() > 0.04 * ROWS
    assert np.all(got.status[degenerate] == 3) and np.all(ref.status[degenerate] == 3)
    assert np.all(np.isnan(got.chi2[degenerate]))
    # The oracle against itself with Kahan sums, and the GPU against the oracle.
    self_agree = np.mean(ref.status == other.status)
    self_both = (ref.status == 0) & (other.status == 0)
    self_within = fp32_gate_rows(other.params[self_both], other.chi2[self_both],
                                 ref.params[self_both], ref.chi2[self_both]).mean()
    agree = np.mean(got.status == ref.status)
    both = (got.status == 0) & (ref.status == 0)
    within = fp32_gate_rows(got.params[both], got.chi2[both], ref.params[both], ref.chi2[both])
    spacing = float(np.min(np.diff(x)))
    for i in np.nonzero(got.status != ref.status)[0]:
        print(f"row {i}: gpu {STATUS[got.status[i]]} (it {got.iterations[i]}, width*2/spacing "
              f"{width_ratio(got.params[i:i + 1], spacing)[0]:.4g}) vs FP32 oracle "
              f"{STATUS[ref.status[i]]} (it {ref.iterations[i]}; {describe(d['stops32'][i])})")
    print(f"doublets numerics={numerics}: status agreement with the FP32 oracle {agree:.5f} "
          f"(oracle plain vs Kahan sums {self_agree:.5f}); converged rows within the FP32 gate "
          f"{within.mean():.4f} of {both.sum()} (oracle vs itself {self_within:.4f}); "
          f"GPU statuses {np.bincount(got.status, minlength=6).tolist()}, oracle "
          f"{np.bincount(ref.status, minlength=6).tolist()}")
    assert agree >= self_agree - 0.005
    assert within.mean() >= self_within - 0.01
    assert np.all(got.iterations <= 100)


@pytest.mark.parametrize("numerics", [0, 5])
def test_config3_doublet_uncertainties(hip, oracle_port, doublets, numerics):
    """tests/uncertainty_checks.hpp:104-190 of the reference: the covariance at
    the GPU's own accepted point against the FP64 oracle's at the same point;
    the fixed intercept and slope (equal bounds) have exactly zero rows,
    columns and standard errors."""
    d = doublets
    x, y, bounds = d["x"], d["y"], d["bounds"]
    got = hip.fit(x, y, "two_lorentzian", None, bounds, "hip", numerics=numerics,
                  uncertainties=True)
    plain = hip.fit(x, y, "two_lorentzian", None, bounds, "hip", numerics=numerics)
    for a, b in zip(plain, got[:4]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), "uncertainties changed the fit"
    ok = got.uncertainty_status == 0
    assert np.all(got.covariance[ok][:, 6:, :] == 0) and np.all(got.covariance[ok][:, :, 6:] == 0)
    assert np.all(got.stderr[ok][:, 6:] == 0)
    assert np.all(got.uncertainty_status[got.status != 0] == 1)
    rows = np.nonzero(got.status == 0)[0]
    outside, status_differs = [], []
    for i in rows:
        cov, se, us = oracle_port.uncertainty("two_lorentzian", x, got.params[i],
                                              int(got.status[i]), float(got.chi2[i]), bounds,
                                              dtype=np.float64)
        if us != got.uncertainty_status[i]:
            status_differs.append(i)
            print(f"row {i}: uncertainty status gpu {got.uncertainty_status[i]} vs FP64 {us}")
        elif us == 0 and not np.allclose(got.covariance[i], cov, rtol=2e-4, atol=1e-9):
            outside.append(i)
            worst = np.max(np.abs(got.covariance[i] - cov) / (1e-9 + np.abs(cov)))
            print(f"row {i}: covariance off by {worst:.3g} relative (stderr gpu "
                  f"{got.stderr[i][:6]} vs {se[:6]})")
    n_ok = int(np.sum(got.uncertainty_status[rows] == 0))
    print(f"doublet uncertainties numerics={numerics}: {n_ok} available of {len(rows)} converged; "
          f"{len(outside)} beyond rtol 2e-4 of the FP64 covariance; "
          f"{len(status_differs)} uncertainty statuses differ")
    assert not status_differs
    assert len(outside) <= 0.005 * max(1, n_ok)


def test_config3_doublets_vs_fp64(hip, doublets):
    d = doublets
    x, y, bounds = d["x"], d["y"], d["bounds"]
    got = hip.fit(x, y, "two_lorentzian", None, bounds, "hip")
    r64, r32, r64t = d["f64"], d["f32"], d["f64_at_f32_tol"]
    spacing = float(np.min(np.diff(x)))
    agree = np.mean(got.status == r64.status)
    ref_agree = np.mean(r32.status == r64.status)
    classes = Counter((STATUS[a], STATUS[b]) for a, b in
                      zip(got.status[got.status != r64.status], r64.status[got.status != r64.status]))
    for i in np.nonzero(got.status != r64.status)[0]:
        print(f"row {i}: gpu {STATUS[got.status[i]]} (it {got.iterations[i]}, width*2/spacing "
              f"{width_ratio(got.params[i:i + 1], spacing)[0]:.4g}; FP32 oracle "
              f"{describe(d['stops32'][i])}) vs FP64 {STATUS[r64.status[i]]} (it "
              f"{r64.iterations[i]}; {describe(d['stops64'][i])})")
    both = (got.status == 0) & (r64.status == 0)
    amp, width, centre, chi = fp64_errors(got, r64, both)
    print(f"doublets vs FP64: status agreement {agree:.5f} (the FP32 oracle: {ref_agree:.5f}; "
          f"the GPU vs FP64 at the FP32 tolerances: {np.mean(got.status == r64t.status):.5f}); "
          f"disagreement classes (gpu, fp64) {dict(classes)}; converged in both {both.sum()}: "
          f"max relative error amplitude {amp.max():.3g}, width {width.max():.3g}, centre/width "
          f"{centre.max():.3g}, chi2 {chi.max():.3g}; median amplitude {np.median(amp):.3g}")
    assert agree >= ref_agree - 0.005
