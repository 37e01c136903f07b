"""Write the golden vectors under tests/golden/ from the reference's own CPU
headers, compiled in place into oracle/_ref/libbfref.so by oracle/Makefile.

Run in a container that has the reference checkout:
    make -C oracle && python tests/golden/make_golden.py
Inputs are regenerated by tests/golden/datasets.py; each fixture stores the
SHA-256 of its inputs so a drifting generator is caught instead of silently
compared. Only outputs and those digests are stored.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import datasets  # noqa: E402
from oracle import oracle  # noqa: E402

SETS = [(32, 257), (129, 257), (1024, 65)]


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return np.frombuffer(h.digest(), np.uint8)


def main():
    ref = oracle.load("reference")
    out = {}
    # README quickstart, float32 and float64, p0=None.
    for name, dtype in (("f32", np.float32), ("f64", np.float64)):
        x, y, bounds = datasets.config1(dtype)
        r = ref.fit(x, y, "lorentzian", None, bounds, threads=4)
        g = ref.guess(x, y, "lorentzian")
        out.update({f"c1_{name}_params": r.params, f"c1_{name}_status": r.status,
                    f"c1_{name}_iterations": r.iterations, f"c1_{name}_chi2": r.chi2,
                    f"c1_{name}_guess": g[0], f"c1_{name}_baseline": g[1],
                    f"c1_{name}_gstatus": g[2], f"c1_{name}_digest": digest(x, y)})
    np.savez_compressed(os.path.join(HERE, "golden_config1.npz"), **out)

    out = {}
    for model in datasets.MODELS:
        for m, n in SETS:
            x, y, p0, bounds, _ = datasets.comparison_set(model, n, m)
            key = f"{model}_{m}"
            out[f"{key}_digest"] = digest(x, y, p0)
            for name, dtype in (("f32", np.float32), ("f64", np.float64)):
                xs, ys = x.astype(dtype), y.astype(dtype)
                given = ref.fit(xs, ys, model, p0, bounds, uncertainties=True, threads=4)
                guessed = ref.fit(xs, ys, model, None, bounds, threads=4)
                comp = ref.fit(xs, ys, model, p0, bounds, threads=4,
                               options=oracle.make_options(compensated=True, scaled=True))
                g = ref.guess(xs, ys, model)
                for tag, r in (("given", given), ("guessed", guessed), ("comp", comp)):
                    for field in ("params", "status", "iterations", "chi2"):
                        out[f"{key}_{name}_{tag}_{field}"] = getattr(r, field)
                out[f"{key}_{name}_given_cov"] = given.covariance
                out[f"{key}_{name}_given_stderr"] = given.stderr
                out[f"{key}_{name}_given_ustatus"] = given.uncertainty_status
                out[f"{key}_{name}_guess"] = g[0]
                out[f"{key}_{name}_baseline"] = g[1]
                out[f"{key}_{name}_gstatus"] = g[2]
    np.savez_compressed(os.path.join(HERE, "golden_sets.npz"), **out)

    out = {}
    x, y, p0 = datasets.isolation_fixture()
    for flags in range(4):
        opts = oracle.make_options(compensated=flags & 1, scaled=flags & 2)
        r = ref.fit(x, y, "lorentzian", p0, None, options=opts)
        out[f"iso_{flags}_status"] = r.status
        out[f"iso_{flags}_iterations"] = r.iterations
        out[f"iso_{flags}_chi2"] = r.chi2
        out[f"iso_{flags}_params"] = r.params
    for model in datasets.MODELS:
        xs, inits, _ = datasets.pathology_fixture(model)
        m = xs.size
        y = np.zeros((9, m), np.float32)
        y[1] = 3
        for i in range(2, 9):
            y[i] = [ref.model(model, float(v), inits[i], np.float32)[0] for v in xs]
        y[2, 17], y[3, 17], y[4, 17] = np.nan, np.inf, -np.inf
        out[f"path_{model}_y"] = y
        for flags in range(4):
            opts = oracle.make_options(compensated=flags & 1, scaled=flags & 2)
            r = ref.fit(xs, y, model, inits, None, options=opts)
            for field in ("params", "status", "iterations", "chi2"):
                out[f"path_{model}_{flags}_{field}"] = getattr(r, field)
    for name, dtype in (("f32", np.float32), ("f64", np.float64)):
        x, y, p0 = datasets.api_isolation_fixture(dtype)
        r = ref.fit(x, y, "lorentzian", p0, None)
        out[f"api_{name}_status"] = r.status
        out[f"api_{name}_params"] = r.params
    np.savez_compressed(os.path.join(HERE, "golden_failures.npz"), **out)
    for f in ("golden_config1.npz", "golden_sets.npz", "golden_failures.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
