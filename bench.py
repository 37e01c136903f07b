"""Benchmark of the MI355X batched bounded LM fitter on the metric workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--rows N] [--no-cpu]

A step is one call of the HIP path over the whole batch already resident in
HBM: grid prologue, data-only guess, LM with covariance off. The workload is
pseudo-Voigt P=4, N=1,048,576 spectra, M=1024 float32 points on
x=linspace(-10,10), amplitude U[0.5,5], centre U[-3,3], FWHM U[0.2,2],
mixing U[0,1], Gaussian noise σ = 1 % of amplitude, seed 7, p0=None,
bounds [0,-10,1e-4,0]–[inf,10,10,1] (tests/golden/datasets.py:config2).
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

METRIC = "fits/sec pseudo-Voigt M=1024 + p50/p99 latency; served vs plain, 1 GPU"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# Algorithmic work per point and pass for pseudo-Voigt, with the counting
# rules of the survey: linearisation 54 flop, cost 16 flop, guess 8 flop.
F_LIN, F_COST, F_GUESS = 54, 16, 8


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=1 << 16)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: fits are independent, so a multi-GPU run is "
              "N copies of this one; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import datasets
    import batchfit_amd
    from batchfit_amd import _native

    lib = _native.library()
    if not lib.bf_hip_available():
        raise SystemExit("no usable gfx950 device")
    n, m = args.rows, 1024
    t0 = time.perf_counter()
    x, y, bounds, _ = datasets.config2(n=1 << 20, rows=n)
    log(f"generated {n}x{m} float32 spectra in {time.perf_counter() - t0:.1f}s")

    err = _native.error_buffer()

    def device(nbytes, host=None):
        ptr = ctypes.c_void_p()
        _native.check(lib.bf_hip_malloc(ctypes.byref(ptr), nbytes, err, len(err)), err)
        if host is not None:
            _native.check(lib.bf_hip_memcpy(ptr, host.ctypes.data, nbytes, 1, err, len(err)), err)
        return ptr

    dx, dy = device(x.nbytes, x), device(y.nbytes, y)
    dparams, dstatus = device(n * 16), device(n * 4)
    diters, dchi2 = device(n * 4), device(n * 4)

    def config(numerics):
        cfg = _native.default_config()
        cfg.fits, cfg.points, cfg.numerics = n, m, numerics
        for k in range(4):
            cfg.lower[k], cfg.upper[k] = bounds[0][k], bounds[1][k]
        return cfg

    def run(cfg, launch, steps, warmup):
        lib.bf_hip_set_timing(1)
        for _ in range(warmup):
            _native.check(lib.bf_hip_fit_device(2, ctypes.byref(cfg), ctypes.byref(launch), dx,
                                                dy, None, dparams, dstatus, diters, dchi2, None,
                                                None, None, err, len(err)), err)
        _native.check(lib.bf_hip_synchronize(err, len(err)), err)
        kernel_ms, step_ms = [], []
        t_start = time.perf_counter()
        for _ in range(steps):
            t = time.perf_counter()
            _native.check(lib.bf_hip_fit_device(2, ctypes.byref(cfg), ctypes.byref(launch), dx,
                                                dy, None, dparams, dstatus, diters, dchi2, None,
                                                None, None, err, len(err)), err)
            times = (ctypes.c_double * 4)()
            lib.bf_hip_last_timing(times, 4)  # waits for this step's last event
            step_ms.append((time.perf_counter() - t) * 1e3)
            kernel_ms.append(list(times))
        _native.check(lib.bf_hip_synchronize(err, len(err)), err)
        elapsed = time.perf_counter() - t_start
        return elapsed, np.array(kernel_ms), np.array(step_ms)

    def outputs():
        params = np.empty((n, 4), np.float32)
        status = np.empty(n, np.int32)
        iterations = np.empty(n, np.uint32)
        chi2 = np.empty(n, np.float32)
        for host, dev in ((params, dparams), (status, dstatus), (iterations, diters),
                          (chi2, dchi2)):
            _native.check(lib.bf_hip_memcpy(host.ctypes.data, dev, host.nbytes, 2, err,
                                            len(err)), err)
        return params, status, iterations, chi2

    # Production form: compensated FP32 partials, FP64 combine/solve, tuned mapping.
    prod = _native.Launch()
    elapsed, kms, sms = run(config(batchfit_amd.NUMERICS_DEFAULT), prod, args.steps, args.warmup)
    params, status, iterations, chi2 = outputs()
    value = n * args.steps / elapsed
    lm_ms = float(kms[:, 3].mean())
    passes = float(np.sum(iterations.astype(np.float64) + 1))
    flops_lm = m * passes * (F_LIN + F_COST)
    achieved = flops_lm / (lm_ms * 1e-3) / 1e12
    traffic = read_traffic()
    log(f"production: {value:,.0f} fits/s, step {elapsed / args.steps * 1e3:.2f} ms, "
        f"kernels ms {np.round(kms.mean(axis=0), 3).tolist()}, mean iterations "
        f"{iterations.mean():.3f}, status counts {np.bincount(status, minlength=6).tolist()}")

    # Plain single-shape launch: FP32 everywhere, lane per fit, point-major y
    # (device transpose), the reference GPU kernel's organisation.
    plain = _native.Launch()
    plain.mapping, plain.point_major = _native.MAP_LANE, 1
    p_elapsed, p_kms, _ = run(config(0), plain, max(1, args.steps // 2), 1)
    p_value = n * max(1, args.steps // 2) / p_elapsed
    log(f"plain: {p_value:,.0f} fits/s, kernels ms {np.round(p_kms.mean(axis=0), 3).tolist()}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(x, y[: args.cpu_rows], bounds)

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded NumPy PCG64, resident in HBM before timing)",
        "config": {
            "workload": "pseudo_voigt P=4 N=1048576 M=1024 float32 x=linspace(-10,10) "
                        "p0=None (device guess) seed 7",
            "fits": n, "points": m, "numerics": "compensated FP32 partials + FP64 solve",
            "mapping": "64 lanes per fit, y in registers",
        },
        "latency_ms": {"p50": float(np.percentile(sms, 50)), "p99": float(np.percentile(sms, 99)),
                       "what": "whole-batch step latency"},
        "kernel_ms": {"grid": float(kms[:, 0].mean()), "guess": float(kms[:, 1].mean()),
                      "lm": lm_ms},
        "plain": {"value": p_value, "unit": "fits/s", "numerics": "FP32",
                  "mapping": "lane per fit, point-major y (device transpose)",
                  "ratio": value / p_value},
        "roofline": {
            "bound": "mfma", "engine": "VALU (FP32; MFMA unused)",
            "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
            "kernel": "lm_kernel<pseudo_voigt, 64 lanes, 16-point tile, compensated, FP64>",
            "algorithmic_flop_per_launch": flops_lm,
            "algorithmic_bytes_per_launch": n * (4 * m + 16 + 12),
            "hbm_achieved_GBs": n * (4 * m + 16 + 12) / (lm_ms * 1e-3) / 1e9,
            "hbm_peak_GBs": PEAK_HBM_GBS,
        },
        "outcome": {"status_counts": np.bincount(status, minlength=6).tolist(),
                    "mean_iterations": float(iterations.mean())},
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    for ptr in (dx, dy, dparams, dstatus, diters, dchi2):
        lib.bf_hip_free(ptr)


def read_traffic():
    """HBM bytes per LM launch from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "lm_pseudo_voigt_m1024_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(x, y, bounds):
    """The reference's own CPU path (oracle/_ref, cpu::fit_batch over its
    ThreadPool) on the first rows of the same workload, float32, p0=None."""
    from oracle import oracle
    kind = "reference"
    try:
        lib = oracle.load("reference")
    except FileNotFoundError:
        kind, lib = "port", oracle.load("port")
    threads = min(16, os.cpu_count() or 1)
    lib.fit(x, y[:256], "pseudo_voigt", None, bounds, threads=threads)  # warm-up
    times = []
    for _ in range(3):
        t = time.perf_counter()
        lib.fit(x, y, "pseudo_voigt", None, bounds, threads=threads)
        times.append(time.perf_counter() - t)
    value = y.shape[0] / float(np.median(times))
    log(f"cpu baseline ({kind}, {threads} threads): {value:,.0f} fits/s")
    return {"value": value, "unit": "fits/s", "cores": threads, "kind": kind,
            "sample": f"first {y.shape[0]} rows of the same workload, float32, p0=None, "
                      "median of 3 after a 256-fit warm-up"}


if __name__ == "__main__":
    main()
