"""CPU checks of the C ABI and the Python boundary: the library loads, exports
every symbol include/batchfit_hip.h declares, keeps the 108-byte Config
layout, and rejects bad arguments before any device work — with the
reference's exception types (python/tests/test_api.py:78-101)."""
import ctypes
import os
import re

import numpy as np
import pytest

import batchfit_amd
from batchfit_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "batchfit_hip.h")).read()
    return sorted(set(re.findall(r"\b(bf_(?:hip|serve|npy)_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.library()
    names = declared_symbols()
    assert names, "no declarations parsed"
    assert sorted(_native.EXPORTS) == names
    for name in names:
        assert hasattr(lib, name), name
    assert lib.bf_hip_abi_version() == 1
    assert lib.bf_hip_available() in (0, 1)


def test_config_layout_and_defaults():
    assert ctypes.sizeof(_native.Config) == 108
    offsets = {name: getattr(_native.Config, name).offset for name, _ in _native.Config._fields_}
    # include/batchfit/metal/types.hpp:7-15 of the reference: 4-byte scalars in order.
    assert offsets["lower"] == 28 and offsets["upper"] == 60 and offsets["numerics"] == 92
    assert offsets["width_factor"] == 104
    cfg = _native.default_config()
    assert cfg.max_iterations == 100
    assert cfg.initial_lambda == pytest.approx(1e-3)
    assert (cfg.gradient_tolerance, cfg.step_tolerance, cfg.cost_tolerance) == pytest.approx(
        (1e-5, 1e-6, 1e-6))
    assert all(np.isinf(cfg.lower[k]) and cfg.lower[k] < 0 for k in range(8))
    assert cfg.numerics == 0


def test_abi_rejects_bad_arguments_without_a_device():
    lib = _native.library()
    cfg = _native.default_config()
    err = _native.error_buffer()
    assert lib.bf_hip_fit(7, ctypes.byref(cfg), None, None, None, None, None, None, None, None,
                          None, None, None, err, len(err)) == _native.BF_INVALID_ARGUMENT
    assert b"unknown fit model" in err.value
    cfg.points = 0
    assert lib.bf_hip_fit(1, ctypes.byref(cfg), None, None, None, None, None, None, None, None,
                          None, None, None, err, len(err)) == _native.BF_INVALID_ARGUMENT
    assert b"M>0" in err.value
    assert lib.bf_hip_fit(1, None, None, None, None, None, None, None, None, None, None, None,
                          None, err, len(err)) == _native.BF_INVALID_ARGUMENT


@pytest.mark.parametrize("change,error", [
    ({"model": "missing"}, ValueError), ({"device": "gpu"}, ValueError),
    ({"device": "cpu"}, ValueError), ({"p0": [1, 2]}, ValueError),
    ({"bounds": ([0, 0], [1, 1])}, ValueError), ({"bounds": (3, 2)}, ValueError),
    ({"bounds": (np.nan, np.inf)}, ValueError), ({"bounds": (-np.inf, 1e300)}, ValueError),
    ({"mapping": "warp"}, ValueError)])
def test_invalid_options(change, error):
    x = np.linspace(-4, 4, 129)
    y = np.tile(2 / (1 + ((x - .2) / .7) ** 2), (8, 1))
    with pytest.raises(error):
        batchfit_amd.fit(x, y, **{"device": "hip", **change})


@pytest.mark.parametrize("x,y,error", [
    (np.zeros((2, 2)), np.zeros((2, 2)), ValueError), ([], [], ValueError),
    ([1, 2], [[1, 2, 3]], ValueError), ([1j, 2j], [[1, 2]], TypeError),
    ([1, 2], [["a", "b"]], TypeError)])
def test_invalid_arrays(x, y, error):
    with pytest.raises(error):
        batchfit_amd.fit(x, y, device="hip")


def test_uncertainty_flag_is_explicit_boolean():
    with pytest.raises(TypeError, match="boolean"):
        batchfit_amd.fit([0, 1, 2], [1, 2, 1], uncertainties="yes")


def test_no_oracle_in_product_path():
    """The product package never imports the CPU oracle."""
    pkg = os.path.join(ROOT, "batchfit_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in text.lower(), f


def test_server_and_plans_need_a_device():
    """Without a usable gfx950 device the server and ragged plans refuse to start
    (RuntimeError), and the serve entries reject null handles."""
    if batchfit_amd.hip_available():
        pytest.skip("a device is present")
    from batchfit_amd.ragged import fit_ragged
    from batchfit_amd.serve import Server
    with pytest.raises(RuntimeError):
        Server()
    x = np.linspace(-1, 1, 32, dtype=np.float32)
    with pytest.raises(RuntimeError):
        fit_ragged([(x, np.ones((2, 32), np.float32), "gaussian", None, None)])
    lib = _native.library()
    import batchfit_amd.serve as serve
    serve._setup(lib)
    ticket = ctypes.c_uint64()
    err = _native.error_buffer()
    cfg = _native.default_config()
    assert lib.bf_serve_submit(None, 0, ctypes.byref(cfg), 0, *([None] * 10), ctypes.byref(ticket),
                               err, len(err)) == _native.BF_INVALID_ARGUMENT
    assert lib.bf_serve_wait(None, 1, 0.0, None) == _native.BF_INVALID_ARGUMENT
    assert lib.bf_serve_destroy(None) == _native.BF_OK


def test_plan_follows_the_tuned_table_without_a_device():
    """bf_hip_plan is host logic: the mapping a launch would use."""
    import re
    from batchfit_amd import plan
    header = open(os.path.join(ROOT, "batchfit_amd", "csrc", "bf_tuned.hpp")).read()
    rows = re.findall(r"\{(\{.*\})\},\s*//\s*(\w+)", header)
    table = {model: [(int(m.rstrip("u"), 0), int(g), int(st)) for m, g, st in
                     re.findall(r"\{(0x[0-9a-f]+u|\d+),\s*(\d+),\s*([01])\}", body)]
             for body, model in rows}
    for model, entries in table.items():
        for m in (5, 64, 100, 256, 300, 512, 1000, 1024, 2048, 4096, 9000):
            bound, g, streamed = next(e for e in entries if m <= e[0])
            per_lane = -(-m // g)
            tile = max(2, 1 << (per_lane - 1).bit_length())
            expect = (g, 0 if streamed or tile > 32 else tile)
            assert plan(model, m) == expect, (model, m)
    assert plan("pseudo_voigt", 1024, mapping="group64") == (64, 16)
    assert plan("pseudo_voigt", 1024, mapping="group64_streamed") == (64, 0)
    assert plan("gaussian", 1500, mapping="group16") == (16, 0)
    assert plan("lorentzian", 100, mapping="lane") == (1, 0)
    # Workgroups of 2 or 4 waves: tiles of 8 or 32 points per lane, else streamed.
    assert plan("pseudo_voigt", 1024, mapping="group256") == (256, 8)
    assert plan("pseudo_voigt", 4096, mapping="workgroup") == (256, 32)
    assert plan("pseudo_voigt", 9000, mapping="group256") == (256, 0)
    assert plan("two_lorentzian", 2048, mapping="group128") == (128, 32)
    assert plan("gaussian", 100, mapping="group128_streamed") == (128, 0)
    # The latency mapping: small batches take wide groups, large ones the table.
    assert plan("pseudo_voigt", 1024, mapping="latency", fits=256) == (64, 16)
    assert plan("pseudo_voigt", 4096, mapping="latency", fits=256) == (128, 32)
    assert plan("pseudo_voigt", 8192, mapping="latency", fits=32) == (256, 32)
    assert plan("pseudo_voigt", 20000, mapping="latency", fits=32) == (256, 0)
    assert plan("pseudo_voigt", 1024, mapping="latency", fits=1 << 20) == plan("pseudo_voigt", 1024)
    assert plan("pseudo_voigt", 4096, mapping="latency", fits=8192) == plan("pseudo_voigt", 4096)
    with pytest.raises(ValueError, match="numerics"):
        plan("gaussian", 256, numerics=1, mapping="group16")


@pytest.mark.parametrize("rows,message", [
    ([0, 1, 1], b"more than one fit"), ([0, 0xFFFFFFFF, 2], b"out of range")])
def test_ragged_plan_rejects_bad_output_rows(rows, message):
    """A ragged plan needs one output row per fit, below 0xffffffff: a repeated
    row would make two fits write the same outputs at once, and 0xffffffff
    would wrap the row count used to size the guess buffer. Both are argument
    errors found on the host, before any device work."""
    from batchfit_amd import ragged
    lib = _native.library()
    ragged._setup(lib)
    fits = np.zeros(3, ragged.FIT_DESC)
    fits["y_offset"] = [0, 32, 64]
    fits["points"] = 32
    fits["p0_row"] = ragged.NO_P0
    fits["out_row"] = rows
    grids = np.zeros(1, ragged.GRID_DESC)
    grids["points"] = 32
    cfg = _native.default_config()
    handle = ctypes.c_void_p()
    err = _native.error_buffer()
    rc = lib.bf_hip_ragged_create(ctypes.byref(handle), 3, fits.ctypes.data, 1, grids.ctypes.data,
                                  1, ctypes.cast(ctypes.pointer(cfg), ctypes.c_void_p), 5, 0,
                                  None, err, len(err))
    assert rc == _native.BF_INVALID_ARGUMENT and message in err.value, (rc, err.value)
    assert not handle.value


@pytest.mark.parametrize("dtype", [np.float64, np.int64])
def test_auto_device_keeps_float64_off_the_gpu(dtype):
    """python/batchfit/__init__.py:84-86 of the reference: device="auto" runs a
    float32 batch on the GPU and keeps float64 (and integer) inputs on its
    float64 CPU path. That path is not in this package, so "auto" refuses such
    inputs instead of narrowing them; "hip" is the explicit float32 request."""
    x = np.linspace(-4, 4, 129).astype(dtype)
    y = np.tile(2 / (1 + ((x - .2) / .7) ** 2), (3, 1)).astype(dtype)
    with pytest.raises(ValueError, match="device='hip'"):
        batchfit_amd.fit(x, y)
    with pytest.raises(ValueError, match="device='hip'"):
        batchfit_amd.fit(x.astype(np.float32), y, device="auto")
