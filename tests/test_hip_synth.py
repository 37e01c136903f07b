"""GPU tests of the device generator (bf_hip_synthesize) against the numpy
Philox4x32-10 checker: drawn parameters, Box–Muller noise on every point, and
independence from how a batch is split (fit_offset)."""
import numpy as np
import pytest

import datasets
from philox import philox4x32_10, synth_parameters, unit

pytestmark = pytest.mark.gpu


def expected_y(model, x, params, noise, seed, fit_offset, relative):
    n, m = params.shape[0], x.size
    f = datasets.MODELS[model]
    p = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8}[model]
    clean = np.stack([f(x.astype(np.float64), *params[i, :p].astype(np.float64)) for i in range(n)])
    fit = np.arange(fit_offset, fit_offset + n, dtype=np.uint64)
    j = np.arange(0, m, 2, dtype=np.uint32)
    F, J = np.meshgrid(fit, j, indexing="ij")
    ctr = np.stack([(F & np.uint64(0xFFFFFFFF)).astype(np.uint32),
                    (F >> np.uint64(32)).astype(np.uint32), J, np.full(J.shape, 7, np.uint32)],
                   axis=-1)
    w = philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32))
    u1, u2 = unit(w[..., 0]).astype(np.float64), unit(w[..., 1]).astype(np.float64)
    rad = np.sqrt(-2.0 * np.log(u1))
    z = np.empty((n, m))
    z[:, 0::2] = (rad * np.cos(2 * np.pi * u2))[:, : (m + 1) // 2]
    z[:, 1::2] = (rad * np.sin(2 * np.pi * u2))[:, : m // 2]
    sigma = noise * params[:, :1].astype(np.float64) if relative else noise
    return clean, z, sigma


@pytest.mark.parametrize("model,layout,relative", [
    ("gaussian", "independent", False), ("pseudo_voigt", "independent", True),
    ("two_lorentzian", "doublet", False)])
def test_generator_matches_philox_checker(hip, model, layout, relative):
    from batchfit_amd.ragged import DeviceArray
    from batchfit_amd.synth import synthesize
    n, m, seed, noise = 300, 257, 0x1234_0000_0042, 0.02
    x = np.linspace(-3, 3, m).astype(np.float32)
    lower = [0.5, -1.0, 0.2, 0.1, 0.0, 0.0, -0.1, -0.01]
    upper = [2.0, 1.0, 0.6, 0.9, 0.0, 0.0, 0.1, 0.01]
    if model == "two_lorentzian":
        lower[3], upper[3] = 0.5, 2.0
    dx, dy, dt = DeviceArray.of(x), DeviceArray(n * m * 4), DeviceArray(n * 32)
    synthesize(model, n, dx, m, lower, upper, dy, noise=noise, noise_relative=relative, seed=seed,
               layout=layout, separation=(2.0, 6.0), truth=dt)
    y = dy.download(np.empty((n, m), np.float32))
    truth = dt.download(np.empty((n, 8), np.float32))
    want = synth_parameters(n, lower, upper, seed, 0, 1 if layout == "doublet" else 0)
    p = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8}[model]
    want[:, p:] = 0.0  # the wrapper passes zero ranges beyond the model's parameters
    # lo + (hi - lo)·u may be contracted to one FMA on the device: one rounding apart.
    eps = np.finfo(np.float32).eps
    scale = np.maximum(np.abs(np.array(lower)), np.abs(np.array(upper)))
    tol = 4 * eps * (scale + np.abs(want))
    if layout == "doublet":
        # c2 = c1 + s·w: two more roundings, relative to the operands (c2 can be ~0).
        tol[:, 4] = 8 * eps * (np.abs(want[:, 1]) + 6.0 * np.abs(want[:, 2]) + np.abs(want[:, 4]))
    assert np.all(np.abs(truth - want) <= tol)
    clean, z, sigma = expected_y(model, x, truth, noise, seed, 0, relative)
    got = y.astype(np.float64)
    tol = 2e-5 * (np.abs(clean) + 1.0) + 1e-4 * np.abs(sigma * z) + 1e-6
    assert np.all(np.abs(got - (clean + sigma * z)) <= tol)
    # Split generation (fit_offset) gives the same bytes.
    dy2 = DeviceArray(n * m * 4)
    half = n // 2
    synthesize(model, half, dx, m, lower, upper, dy2, noise=noise, noise_relative=relative,
               seed=seed, layout=layout)
    tail = DeviceArray(n * m * 4)
    synthesize(model, n - half, dx, m, lower, upper, tail, noise=noise, noise_relative=relative,
               seed=seed, fit_offset=half, layout=layout)
    a = dy2.download(np.empty((half, m), np.float32))
    b = tail.download(np.empty((n - half, m), np.float32))
    assert np.array_equal(np.concatenate([a, b]).view(np.uint32), y.view(np.uint32))


def generate(model, n, m, x, lower, upper, mode, noise, seed=0xBA7C, fit_offset=0):
    from batchfit_amd.ragged import DeviceArray
    from batchfit_amd.synth import synthesize
    dx, dy = DeviceArray.of(x), DeviceArray(n * m * 4)
    synthesize(model, n, dx, m, lower, upper, dy, noise=noise, noise_mode=mode, seed=seed,
               fit_offset=fit_offset)
    y = dy.download(np.empty((n, m), np.float32))
    dx.free()
    dy.free()
    return y


def test_snr_gaussian_noise(hip):
    """src/cpu/synthetic.cpp:83-89 of the reference: σ = RMS of the clean row /
    snr, per row; the normals are the generator's Box–Muller pairs."""
    n, m, snr, seed = 200, 257, 20.0, 0xBA7C
    x = np.linspace(-3, 3, m).astype(np.float32)
    lower, upper = [0.5, -1, 0.2, 0.1], [2, 1, 0.6, 0.9]
    clean = generate("pseudo_voigt", n, m, x, lower, upper, "absolute", 0.0, seed).astype(np.float64)
    y = generate("pseudo_voigt", n, m, x, lower, upper, "snr_gaussian", snr, seed)
    params = synth_parameters(n, lower + [0] * 4, upper + [0] * 4, seed)
    _, z, _ = expected_y("pseudo_voigt", x, params, 0.0, seed, 0, False)
    sigma = np.sqrt(np.mean(clean ** 2, axis=1, keepdims=True)) / snr
    want = clean + sigma * z
    # The clean value is one FP32 evaluation of the model (an ulp of it may
    # differ between the two launches), the sum is rounded once to FP32, and
    # the device's Box–Muller angle 2π·u is an FP32 value (up to ~5e-7 rad off,
    # times a radius below 6, in units of σ).
    ulp = lambda v: np.spacing(np.abs(v).astype(np.float32)).astype(np.float64)
    tol = sigma * (1e-5 * np.abs(z) + 4e-6) + 2 * ulp(clean) + 2 * ulp(want)
    assert np.all(np.abs(y - want) <= tol)
    resid = (y - clean) / sigma
    assert abs(resid.mean()) < 0.02 and abs(resid.std() - 1) < 0.02


def test_snr_poisson_noise(hip):
    """src/cpu/synthetic.cpp:90-108 of the reference: gain = snr²·Σf/Σf²,
    y = Poisson(gain·f)/gain. Counts are checked against the numpy checker
    (inversion below a mean of 10, PTRS from 10) on the same Philox streams,
    and as a distribution."""
    from philox import poisson_count
    n, m, snr, seed = 40, 129, 6.0, 0x5EED
    x = np.linspace(-4, 4, m).astype(np.float32)
    lower, upper = [0.5, -1, 0.3], [3, 1, 1.5]
    clean = generate("gaussian", n, m, x, lower, upper, "absolute", 0.0, seed).astype(np.float64)
    y = generate("gaussian", n, m, x, lower, upper, "snr_poisson", snr, seed).astype(np.float64)
    gain = snr * snr * clean.sum(axis=1, keepdims=True) / (clean ** 2).sum(axis=1, keepdims=True)
    mean = gain * clean
    counts = y * gain
    assert np.all(np.abs(counts - np.rint(counts)) <= 1e-3 * np.maximum(1, counts))
    counts = np.rint(counts)
    want = np.array([[poisson_count(mean[i, j], i, j, seed) for j in range(m)] for i in range(n)])
    same = np.mean(counts == want)
    assert same >= 0.999, same
    for lo, hi in ((0.5, 10), (10, np.inf)):
        sel = (mean >= lo) & (mean < hi)
        zs = (counts[sel] - mean[sel]) / np.sqrt(mean[sel])
        assert sel.sum() > 300, (lo, sel.sum())
        assert abs(zs.mean()) < 5 / np.sqrt(sel.sum()) and abs(zs.var() - 1) < 0.15, (lo, zs.var())
    # The same counts when the batch is generated in two pieces.
    a = generate("gaussian", 15, m, x, lower, upper, "snr_poisson", snr, seed)
    b = generate("gaussian", n - 15, m, x, lower, upper, "snr_poisson", snr, seed, fit_offset=15)
    assert np.array_equal(np.concatenate([a, b]).view(np.uint32), y.astype(np.float32).view(np.uint32))


def test_poisson_rejects_negative_signal_and_bad_snr(hip):
    from batchfit_amd.ragged import DeviceArray
    from batchfit_amd.synth import synthesize
    m = 64
    x = np.linspace(-2, 2, m).astype(np.float32)
    # A doublet on a negative intercept: every row has negative values.
    lower = [1, -1, 0.2, 1, 0.5, 0.2, -5, 0]
    upper = [2, 0, 0.4, 2, 1.0, 0.4, -4, 0]
    y = generate("two_lorentzian", 8, m, x, lower, upper, "snr_poisson", 10.0)
    assert np.all(np.isnan(y))
    dx, dy = DeviceArray.of(x), DeviceArray(8 * m * 4)
    for snr in (0.0, -1.0, np.inf):
        with pytest.raises(ValueError, match="snr"):
            synthesize("gaussian", 8, dx, m, [1, 0, 1], [1, 0, 1], dy, noise=snr,
                       noise_mode="snr_gaussian")
