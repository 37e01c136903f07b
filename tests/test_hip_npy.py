"""GPU tests of the NPY streaming pipeline (bf_hip_fit_npy) and its CLI.

The pipeline must give the bytes of the uniform entry on the same rows:
float32 files directly, float64 files after rounding to float32 (narrowed on
the device, as NumPy's astype rounds), for any chunk size and ring depth,
chunks handed over in order."""
import os

import numpy as np
import pytest

import datasets

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                          np.ascontiguousarray(b).view(np.uint8))


def files(tmp_path, x, y, p0=None):
    np.save(tmp_path / "x.npy", x)
    np.save(tmp_path / "y.npy", y)
    if p0 is not None:
        np.save(tmp_path / "p0.npy", p0)
    return str(tmp_path / "x.npy"), str(tmp_path / "y.npy"), (
        None if p0 is None else str(tmp_path / "p0.npy"))


@pytest.mark.parametrize("chunk,depth", [(0, 3), (333, 1), (333, 3), (64, 2)])
def test_float32_file_equals_uniform(hip, tmp_path, chunk, depth):
    from batchfit_amd.npy import fit_npy
    x, y, p0, bounds, _ = datasets.comparison_set("pseudo_voigt", 2000, 257, seed=9)
    x, y = x.astype(np.float32), y.astype(np.float32)
    xp, yp, _ = files(tmp_path, x, y)
    got, stats = fit_npy(xp, yp, "pseudo_voigt", None, bounds, chunk_fits=chunk, in_flight=depth)
    ref = hip.fit(x, y, "pseudo_voigt", None, bounds, "hip")
    for a, b in zip(got, ref):
        assert same(a, b)
    assert stats["fits"] == 2000 and stats["chunks"] == -(-2000 // stats["chunk_fits"])
    assert 1 <= stats["max_in_flight"] <= depth


def test_float64_files_with_p0_and_uncertainties(hip, tmp_path):
    from batchfit_amd.npy import fit_npy
    x, y, p0, bounds, _ = datasets.comparison_set("two_lorentzian", 700, 400, seed=4)
    p0 = np.broadcast_to(np.asarray(p0), (y.shape[0], 8))
    x64, y64 = x.astype(np.float64), y.astype(np.float64)
    p64 = np.ascontiguousarray(p0, np.float64)
    xp, yp, pp = files(tmp_path, x64, y64, p64)
    got, stats = fit_npy(xp, yp, "two_lorentzian", pp, bounds, uncertainties=True, chunk_fits=256)
    ref = hip.fit(x64.astype(np.float32), y64.astype(np.float32), "two_lorentzian",
                  p64.astype(np.float32), bounds, "hip", uncertainties=True)
    for a, b in zip(got, ref):
        assert same(a, b)
    assert stats["chunks"] == 3


def test_consumer_order_and_errors(hip, tmp_path):
    from batchfit_amd.npy import fit_npy
    x, y, p0, bounds, _ = datasets.comparison_set("gaussian", 1000, 129, seed=2)
    xp, yp, _ = files(tmp_path, x.astype(np.float32), y.astype(np.float32))
    seen = []
    result, stats = fit_npy(xp, yp, "gaussian", consumer=lambda first, r: seen.append(
        (first, r.params.shape[0])), chunk_fits=300)
    assert result is None
    assert seen == [(0, 300), (300, 300), (600, 300), (900, 100)]

    def boom(first, r):
        raise KeyError("stop")
    with pytest.raises(KeyError):
        fit_npy(xp, yp, "gaussian", consumer=boom, chunk_fits=300)
    np.save(tmp_path / "short.npy", x[:-1].astype(np.float32))
    with pytest.raises(ValueError, match="x\\[M\\], y\\[N,M\\]"):
        fit_npy(str(tmp_path / "short.npy"), yp, "gaussian")
    np.save(tmp_path / "p0bad.npy", np.zeros((1000, 4), np.float32))
    with pytest.raises(ValueError, match="initial"):
        fit_npy(xp, yp, "gaussian", str(tmp_path / "p0bad.npy"))


def test_cli_writes_typed_outputs(hip, tmp_path):
    from batchfit_amd.fit_npy import main
    x, y, p0, bounds, _ = datasets.comparison_set("lorentzian", 500, 129, seed=6)
    xp, yp, _ = files(tmp_path, x.astype(np.float32), y.astype(np.float32))
    out = tmp_path / "out"
    assert main(["--x", xp, "--y", yp, "--output", str(out), "--model", "lorentzian",
                 "--chunk-fits", "128", "--timeline", str(tmp_path / "t.json")]) == 0
    ref = hip.fit(x.astype(np.float32), y.astype(np.float32), "lorentzian", None, None, "hip")
    assert same(np.load(out / "parameters.npy"), ref.params)
    assert np.load(out / "status.npy").dtype == np.uint32
    assert same(np.load(out / "status.npy").astype(np.int32), ref.status)
    assert same(np.load(out / "iterations.npy"), ref.iterations)
    assert same(np.load(out / "chi_squared.npy"), ref.chi2)
    assert main(["--x", xp, "--y", yp, "--output", str(out)]) == 1  # directory exists
    assert main(["--x", xp, "--y", str(tmp_path / "missing.npy"), "--output",
                 str(tmp_path / "o2")]) == 1
    assert not os.path.exists(tmp_path / "o2")
