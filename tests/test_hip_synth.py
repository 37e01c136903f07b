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
