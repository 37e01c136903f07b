"""The no-allocation entry point (bf_hip_fit_device_noalloc) against
bf_hip_fit_device on the same device inputs: the same bytes, over two calls
in a row (the second smaller, so a stale guess row or counter left in the
shared scratch by the first call would show), with and without uncertainties,
from the device guess and from given p0; sizes beyond its scratch are refused."""
import ctypes

import numpy as np
import pytest

import datasets

pytestmark = pytest.mark.gpu

P = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8}


def run(entry, model, x, y, p0, bounds, unc, launch=None):
    from batchfit_amd import _native
    from batchfit_amd.ragged import DeviceArray
    lib = _native.library()
    n, m = y.shape
    p = P[model]
    cfg = _native.default_config()
    cfg.fits, cfg.points, cfg.numerics = n, m, 5
    for k in range(p):
        cfg.lower[k], cfg.upper[k] = float(bounds[0][k]), float(bounds[1][k])
    dx, dy = DeviceArray.of(np.ascontiguousarray(x, np.float32)), DeviceArray.of(
        np.ascontiguousarray(y, np.float32))
    dp0 = None if p0 is None else DeviceArray.of(
        np.ascontiguousarray(np.broadcast_to(p0, (n, p)), np.float32))
    outs = [DeviceArray(n * p * 4), DeviceArray(n * 4), DeviceArray(n * 4), DeviceArray(n * 4)]
    if unc:
        outs += [DeviceArray(n * p * p * 4), DeviceArray(n * p * 4), DeviceArray(n * 4)]
    ptrs = [o.ptr for o in outs] + ([None] * 3 if not unc else [])
    err = _native.error_buffer()
    rc = getattr(lib, entry)(_native.MODELS[model], ctypes.byref(cfg),
                             None if launch is None else ctypes.byref(launch), dx.ptr, dy.ptr,
                             None if dp0 is None else dp0.ptr, *ptrs, err, len(err))
    _native.check(rc, err)
    err2 = _native.error_buffer()
    _native.check(lib.bf_hip_synchronize(err2, len(err2)), err2)
    shapes = [(n, p), (n,), (n,), (n,), (n, p, p), (n, p), (n,)]
    dtypes = [np.float32, np.int32, np.uint32, np.float32, np.float32, np.float32, np.int32]
    got = [o.download(np.empty(s, d)) for o, s, d in zip(outs, shapes, dtypes)]
    for a in outs + [dx, dy] + ([dp0] if dp0 is not None else []):
        a.free()
    return got


def same_bytes(a, b, label):
    assert len(a) == len(b)
    for u, v in zip(a, b):
        assert np.array_equal(u.view(np.uint8), v.view(np.uint8)), label


@pytest.mark.parametrize("unc", [False, True])
def test_noalloc_matches_device_entry_twice(hip, unc):
    x, y, bounds, _ = datasets.config2(rows=12000)
    xs, ys, p0s, bs, _ = datasets.comparison_set("two_lorentzian", 3000, 300)
    calls = [("pseudo_voigt", x, y, None, bounds),               # device guess, 12,000 fits
             ("two_lorentzian", xs, ys[:2500], None, bs),       # smaller, other model
             ("two_lorentzian", xs, ys[:700], p0s, bs)]         # given p0, smaller again
    for model, xx, yy, p0, bb in calls:
        print(f"noalloc vs device entry: {model} N={yy.shape[0]} M={yy.shape[1]} "
              f"p0={'given' if p0 is not None else 'guess'} uncertainties={unc}", flush=True)
        a = run("bf_hip_fit_device_noalloc", model, xx, yy, p0, bb, unc)
        b = run("bf_hip_fit_device", model, xx, yy, p0, bb, unc)
        same_bytes(a, b, f"{model} N={yy.shape[0]}")
        assert np.bincount(a[1], minlength=6)[0] > 0


def test_noalloc_refuses_what_its_scratch_cannot_hold(hip):
    from batchfit_amd import _native
    lib = _native.library()
    cfg = _native.default_config()
    cfg.fits, cfg.points = 1 << 22, 64  # 4,194,304 guess rows of P=4 need 64 MiB
    err = _native.error_buffer()
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lib.bf_hip_fit_device_noalloc(2, ctypes.byref(cfg), None, dummy, dummy, None,
                                       *([dummy] * 4), None, None, None, err, len(err))
    assert rc == _native.BF_INVALID_ARGUMENT and b"at most" in err.value
    launch = _native.Launch()
    launch.mapping, launch.point_major = _native.MAP_LANE, 1
    cfg.fits = 16
    rc = lib.bf_hip_fit_device_noalloc(2, ctypes.byref(cfg), ctypes.byref(launch), dummy, dummy,
                                       None, *([dummy] * 4), None, None, None, err, len(err))
    assert rc == _native.BF_INVALID_ARGUMENT and b"point-major" in err.value
