"""GPU tests of the continuous-batching server (bf_serve_*).

* Same alone and served: every request's bytes equal the uniform entry's on
  the same batch, whatever else was admitted with it, for host requests
  submitted from several threads and for device-resident requests.
* Admission: with an HBM budget of about one request, no two requests are on
  the device together and every request still completes.
* Errors: argument errors raise at submission; a batch's data failures stay in
  its status rows.
"""
import os
import subprocess
import threading

import numpy as np
import pytest

import datasets

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                          np.ascontiguousarray(b).view(np.uint8))


def requests():
    out = []
    for i, model in enumerate(datasets.MODELS):
        for m, n in ((48, 23), (200, 31), (700, 17), (1800, 9), (3000, 4)):
            x, y, p0, bounds, _ = datasets.comparison_set(model, n, m, seed=100 + i * 10 + m)
            out.append((x.astype(np.float32), y.astype(np.float32), model,
                        None if (m + i) % 2 else p0, bounds))
    return out


def test_served_long_spectra_from_guess(hip):
    """A request of doublets longer than the one-wave guess's LDS (M = 16,384)
    without p0 is served from the wide guess: the uniform entry's bytes."""
    from batchfit_amd.serve import Server
    x, y, _, bounds, _ = datasets.comparison_set("two_lorentzian", 5, 16384)
    x, y = x.astype(np.float32), y.astype(np.float32)
    with Server(slots=2) as srv:
        got = srv.submit(x, y, "two_lorentzian", None, bounds).wait(timeout=60)
    ref = hip.fit(x, y, "two_lorentzian", None, bounds, "hip")
    for a, b in zip(got, ref):
        assert same(a, b)


def test_served_results_equal_uniform_fits(hip):
    from batchfit_amd.serve import Server
    batch = requests()
    tickets = [None] * len(batch)

    def submit(srv, idx):
        for i in idx:
            x, y, model, p0, bounds = batch[i]
            tickets[i] = srv.submit(x, y, model, p0, bounds)

    with Server(slots=3) as srv:
        workers = [threading.Thread(target=submit, args=(srv, range(t, len(batch), 4)))
                   for t in range(4)]
        for w in workers:
            w.start()
        for w in workers:
            w.join()
        results = [t.wait(timeout=60) for t in tickets]
        stats = srv.stats()
    assert stats["requests_done"] == len(batch)
    assert stats["fits_done"] == sum(b[1].shape[0] for b in batch)
    for (x, y, model, p0, bounds), got, ticket in zip(batch, results, tickets):
        ref = hip.fit(x, y, model, p0, bounds, "hip")
        for a, b in zip(got, ref):
            assert same(a, b), f"{model} M={x.size}"
        assert 0 <= ticket.timing.service <= ticket.timing.latency


def test_budget_admits_one_request_at_a_time(hip):
    from batchfit_amd.serve import Server
    batch = [b for b in requests() if b[0].size == 700]

    def held(x, y, p0):
        # Bytes a request holds on the device: grid, observations, per-fit
        # descriptor/order/guess/output rows, its config, and p0 rows.
        n, m = y.shape
        return 4 * (m + n * m) + 4 * n * 19 + 32 * n + 160 + (32 * n if p0 is not None else 0)

    sizes = [held(b[0], b[1], b[3]) for b in batch]
    budget = int(max(sizes) * 1.2)
    assert 2 * min(sizes) > budget
    with Server(hbm_budget=budget, slots=3) as srv:
        tickets = [srv.submit(*b) for b in batch]
        results = [t.wait(timeout=60) for t in tickets]
        stats = srv.stats()
    assert stats["max_bytes_in_flight"] <= budget
    assert stats["batches"] == len(batch)
    for (x, y, model, p0, bounds), got in zip(batch, results):
        ref = hip.fit(x, y, model, p0, bounds, "hip")
        assert same(got.params, ref.params) and same(got.status, ref.status)


def test_device_requests_and_uncertainties(hip, oracle_port):
    from batchfit_amd.ragged import DeviceArray
    from batchfit_amd.serve import Server
    x, y, p0, bounds, _ = datasets.comparison_set("pseudo_voigt", 40, 512, seed=5)
    x, y = x.astype(np.float32), y.astype(np.float32)
    y[2, 11] = np.nan
    y[3] = 0.5
    n, m, p = 40, 512, 4
    dx, dy = DeviceArray.of(x), DeviceArray.of(y)
    out = {"params": DeviceArray(n * p * 4), "status": DeviceArray(n * 4),
           "iterations": DeviceArray(n * 4), "chi2": DeviceArray(n * 4)}
    with Server() as srv:
        on_device = srv.submit_device(dx, dy, "pseudo_voigt", n, m, out, None, bounds)
        on_host = srv.submit(x, y, "pseudo_voigt", None, bounds, uncertainties=True)
        on_device.wait(timeout=60)
        host = on_host.wait(timeout=60)
    ref = hip.fit(x, y, "pseudo_voigt", None, bounds, "hip", uncertainties=True)
    for a, b in zip(host, ref):
        assert same(a, b)
    params = out["params"].download(np.empty((n, p), np.float32))
    status = out["status"].download(np.empty(n, np.int32))
    assert same(params, ref.params) and same(status, ref.status)
    # Failure rows (a NaN point, a flat row) keep the oracle's statuses.
    ref_cpu = oracle_port.fit(x, y, "pseudo_voigt", None, bounds, threads=8)
    assert np.array_equal(status, ref_cpu.status) and status[2] == 3


def test_submission_errors(hip):
    from batchfit_amd.serve import Server
    x = np.linspace(-1, 1, 64, dtype=np.float32)
    y = np.ones((3, 64), np.float32)
    with Server() as srv:
        with pytest.raises(ValueError):
            srv.submit(x, y, "voigt")
        with pytest.raises(ValueError):
            srv.submit(x, y[:, :10], "gaussian")
        with pytest.raises(ValueError):
            srv.submit(x, y, "gaussian", None, ([0, 0, 1], [1, 1, 0]))
        empty = srv.submit(x, np.zeros((0, 64), np.float32), "gaussian")
        assert empty.wait(timeout=10).params.shape == (0, 3)
    with Server(hbm_budget=4096) as small:
        with pytest.raises(ValueError, match="budget"):
            small.submit(x, np.ones((100, 64), np.float32), "gaussian")


def test_server_host_code_under_address_sanitizer():
    """The server's host code built with AddressSanitizer and UBSan
    (tests/cpp/Makefile, target asan, part of __graft_entry__.build()) and
    driven by tests/cpp/serve_stress.cpp: 400 host and device requests from
    four threads while a fifth polls the counters, batches capped at 512 fits
    so admission and completion interleave; no sanitizer report, and every
    result equals bf_hip_fit on the same request."""
    exe = os.path.join(os.path.dirname(__file__), "cpp", "_build", "serve_stress_asan")
    assert os.path.exists(exe), "build it with `make -C tests/cpp asan`"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    out = subprocess.run([exe, "400"], capture_output=True, text=True, env=env, timeout=240)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert "mismatches 0" in out.stdout and "Sanitizer" not in out.stderr, out.stderr[-4000:]
