"""Command-line NPY fitter (the reference's tools/fit_npy.cpp, `--device hip`).

    python -m batchfit_amd.fit_npy --x x.npy --y y.npy [--initial p0.npy]
        --output NEW_DIR [--model gaussian|lorentzian|pseudo_voigt|two_lorentzian]
        [--chunk-fits 0] [--in-flight 3] [--uncertainties] [--timeline events.json]

Streams y through the device pipeline and writes typed NPY outputs into a new
directory: parameters.npy [N,P] float32, chi_squared.npy [N] float32,
status.npy [N] uint32, iterations.npy [N] uint32 (and covariance.npy,
stderr.npy, uncertainty_status.npy with --uncertainties). The directory is
removed again if the run fails (tools/fit_npy.cpp:26-64). Prints one line of
pipeline counters.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import sys
import time

import numpy as np


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m batchfit_amd.fit_npy")
    ap.add_argument("--x", required=True)
    ap.add_argument("--y", required=True)
    ap.add_argument("--initial")
    ap.add_argument("--output", required=True)
    ap.add_argument("--model", default="gaussian",
                    choices=("gaussian", "lorentzian", "pseudo_voigt", "two_lorentzian"))
    ap.add_argument("--chunk-fits", type=int, default=0)
    ap.add_argument("--in-flight", type=int, default=3)
    ap.add_argument("--uncertainties", action="store_true")
    ap.add_argument("--timeline")
    args = ap.parse_args(argv)
    if not 1 <= args.in_flight <= 8:
        print("in-flight must be 1..8", file=sys.stderr)
        return 1
    from . import _PARAMETERS
    from .npy import fit_npy, inspect
    try:
        header = inspect(args.y)
        if len(header.shape) != 2:
            raise ValueError("y must have shape [N,M]")
        n, p = header.shape[0], _PARAMETERS[args.model]
        os.mkdir(args.output)  # must not exist yet
    except (OSError, ValueError, RuntimeError) as exc:
        print(exc, file=sys.stderr)
        return 1
    try:
        def out(name, shape, dtype):
            return np.lib.format.open_memmap(os.path.join(args.output, name), mode="w+",
                                             dtype=dtype, shape=shape)
        files = [out("parameters.npy", (n, p), np.float32), out("status.npy", (n,), np.uint32),
                 out("iterations.npy", (n,), np.uint32), out("chi_squared.npy", (n,), np.float32)]
        if args.uncertainties:
            files += [out("covariance.npy", (n, p, p), np.float32),
                      out("stderr.npy", (n, p), np.float32),
                      out("uncertainty_status.npy", (n,), np.uint32)]
        events = []
        begin = time.perf_counter()

        def consume(first, chunk):
            t = time.perf_counter()
            count = chunk.status.shape[0]
            for dst, src in zip(files, chunk):
                dst[first:first + count] = src.astype(dst.dtype, copy=False)
            events.append({"name": f"fits {first}..{first + count}", "ph": "X", "pid": 0,
                           "tid": 0, "ts": (t - begin) * 1e6,
                           "dur": (time.perf_counter() - t) * 1e6})

        _, stats = fit_npy(args.x, args.y, args.model, args.initial,
                           uncertainties=args.uncertainties, chunk_fits=args.chunk_fits,
                           in_flight=args.in_flight, consumer=consume)
        for f in files:
            f.flush()
        del files
        if args.timeline:
            with open(args.timeline, "w") as fh:
                json.dump({"traceEvents": events}, fh)
        wall = time.perf_counter() - begin
        print(f"fits={stats['fits']} chunks={stats['chunks']} chunk_fits={stats['chunk_fits']} "
              f"peak_pending={stats['max_in_flight']} pipeline_seconds={stats['wall_seconds']:.6f} "
              f"wall_seconds={wall:.6f} fits_per_second={stats['fits'] / wall if wall else 0:.1f}")
        return 0
    except BaseException as exc:
        shutil.rmtree(args.output, ignore_errors=True)
        print(exc, file=sys.stderr)
        return 1


if __name__ == "__main__":
    sys.exit(main())
