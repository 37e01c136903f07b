"""Autotuner of the fit-to-hardware mapping (lanes per fit G, by model and M).

    python -m batchfit_amd.autotune [--fits 65536] [--repeat 3] [--write]

For every model and a ladder of point counts M, fits of that shape are
generated in HBM (Philox generator, parameters as in the ragged benchmark
workload), started from the device guess as in production (so the spread
of iteration counts between the fits that share a wave is measured too), and
solved by the LM kernel with each candidate mapping — 8, 16, 32 or 64 lanes
per fit, with the points either in a register tile (the smallest power of two
≥ M/G, up to 32 per lane) or read from memory on every pass, or 4 lanes per
fit streamed — in the
production numerics (compensated partials, FP64 solve). A fit's result
depends on G and M only, so the table changes speed, not results. The LM
kernel time is read from the library's events; the fastest mapping wins
each (model, M), and a dynamic programme over the ladder trades a slower
mapping against a switch between mappings (a switch costs 10 %). The
resulting ranges are,
with --write, emitted as batchfit_amd/csrc/bf_tuned.hpp (the table both the
uniform and the ragged paths use), followed by a rebuild. The measured table
is printed as JSON.

The reference picks among its kernel variants by a measured policy for its
Apple GPU (include/batchfit/metal/batch.hpp:21-50); this is the same idea for
gfx950, measured rather than hand-written.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys

import numpy as np

from . import _native

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = ("gaussian", "lorentzian", "pseudo_voigt", "two_lorentzian")
LADDER = (32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096)
# (lanes per fit, streamed): points in a register tile, or read from memory
# on every pass (4-lane groups always stream).
CANDIDATES = ((8, 0), (16, 0), (32, 0), (64, 0), (4, 1), (8, 1), (16, 1), (32, 1), (64, 1))
RANGES = {  # parameters drawn per fit (the ragged benchmark workload's ranges)
    "gaussian": ([0.5, -3, 0.085], [5, 3, 0.85]),
    "lorentzian": ([0.5, -3, 0.1], [5, 3, 1.0]),
    "pseudo_voigt": ([0.5, -3, 0.2, 0.0], [5, 3, 2.0, 1.0]),
    "two_lorentzian": ([0.5, -5, 0.2, 0.5, 0, 0, 0, 0], [5, 5, 1.5, 5, 0, 0, 0, 0]),
}
NOISE = {"gaussian": (0.01, True), "lorentzian": (0.01, True), "pseudo_voigt": (0.01, True),
         "two_lorentzian": (0.01, False)}  # σ, relative to the amplitude or absolute
P = {"gaussian": 3, "lorentzian": 3, "pseudo_voigt": 4, "two_lorentzian": 8}


def tile(g: int, m: int) -> int:
    r = -(-m // g)
    t = 2
    while t < r:
        t *= 2
    return t if t <= 32 else 0


SWITCH = 0.10  # relative slowdown one change of mapping along the ladder must save


def label(g: int, streamed: int) -> str:
    return f"{g}s" if streamed else str(g)


def unlabel(key) -> tuple:
    key = str(key)
    return (int(key[:-1]), 1) if key.endswith("s") else (int(key), 0)


def measure(fits: int, repeat: int) -> dict:
    from .ragged import DeviceArray
    from .synth import synthesize
    lib = _native.library()
    if not lib.bf_hip_available():
        raise SystemExit("no usable gfx950 device")
    err = _native.error_buffer()
    lib.bf_hip_set_timing(1)
    table = {}
    for model in MODELS:
        p = P[model]
        lo, hi = RANGES[model]
        for m in LADDER:
            x = np.linspace(-10, 10, m).astype(np.float32)
            dx, dy = DeviceArray.of(x), DeviceArray(fits * m * 4)
            noise, relative = NOISE[model]
            synthesize(model, fits, dx, m, lo, hi, dy, noise=noise, noise_relative=relative,
                       seed=41, layout="doublet" if model == "two_lorentzian" else "independent",
                       separation=(0.5, 3.0))
            outs = [DeviceArray(fits * p * 4), DeviceArray(fits * 4), DeviceArray(fits * 4),
                    DeviceArray(fits * 4)]
            cfg = _native.default_config()
            cfg.fits, cfg.points, cfg.numerics = fits, m, 5
            times = {}
            for g, streamed in CANDIDATES:
                if not streamed and tile(g, m) == 0:
                    continue  # beyond the largest register tile
                launch = _native.Launch()
                launch.mapping, launch.group_lanes = _native.MAP_GROUP, g
                launch.streamed = streamed
                best = float("inf")
                for _ in range(repeat + 1):
                    _native.check(lib.bf_hip_fit_device(
                        _native.MODELS[model], ctypes.byref(cfg), ctypes.byref(launch), dx.ptr,
                        dy.ptr, None, *[o.ptr for o in outs], None, None, None, err,
                        len(err)), err)
                    t = (ctypes.c_double * 4)()
                    lib.bf_hip_last_timing(t, 4)
                    best = min(best, t[3])
                times[label(g, streamed)] = best
            winner = min(times, key=times.get)
            table[(model, m)] = (winner, times)
            print(f"{model:15s} M={m:5d} " +
                  " ".join(f"G{k}={times[k]:.3f}ms" for k in times) + f" -> {winner}",
                  file=sys.stderr, flush=True)
            for a in [dx, dy] + outs:
                a.free()
    return table


def choose(rows: list) -> list:
    """One mapping per ladder point: the sequence that minimises the summed
    relative slowdown (time / fastest − 1) plus SWITCH for every change of
    mapping between neighbours, so that measurement noise does not split the
    table into many small ranges (dynamic programming over the ladder)."""
    labels = sorted({k for _, times in rows for k in times})
    inf = float("inf")

    def cost(times, k):
        return times[k] / min(times.values()) - 1.0 if k in times else inf

    best = {k: cost(rows[0][1], k) for k in labels}
    back = []
    for _, times in rows[1:]:
        prev = min(best, key=best.get)
        step, links = {}, {}
        for k in labels:
            stay, move = best[k], best[prev] + SWITCH
            links[k] = k if stay <= move else prev
            step[k] = min(stay, move) + cost(times, k)
        back.append(links)
        best = step
    k = min(best, key=best.get)
    out = [k]
    for links in reversed(back):
        k = links[k]
        out.append(k)
    return out[::-1]


def ranges(table: dict) -> dict:
    """Per model: list of (largest M, lanes, streamed) covering every M.
    Mappings are chosen along the ladder by choose(); between two mappings
    the boundary is the midpoint of their M, but a register-tile mapping with
    fewer than 64 lanes never extends beyond what its largest tile holds (32
    points per lane). The last range is open."""
    out = {}
    for model in MODELS:
        rows = [(m, table[(model, m)][1] or {table[(model, m)][0]: 1.0}) for m in LADDER]
        picks = choose(rows)
        pts = [(m, unlabel(k)) for (m, _), k in zip(rows, picks)]
        merged = []
        for i, (m, (g, st)) in enumerate(pts):
            cap = 0xFFFFFFFF if (st or g == 64) else 32 * g
            upper = min((m + pts[i + 1][0]) // 2, cap) if i + 1 < len(pts) else cap
            if merged and merged[-1][1:] == (g, st):
                merged[-1] = (upper, g, st)
            else:
                merged.append((upper, g, st))
        if merged[-1][0] != 0xFFFFFFFF:
            merged.append((0xFFFFFFFF, 64, 0))
        out[model] = merged
    return out


def write_header(rng: dict, path: str) -> None:
    width = max(len(v) for v in rng.values())
    lines = [
        "// bf_tuned.hpp — generated by `python -m batchfit_amd.autotune --write`: lanes per",
        "// fit by (model, M), measured on gfx950 (LM kernel time of the production numerics",
        "// from device-guess starts; see batchfit_amd/autotune.py). Rows: gaussian, lorentzian,",
        "// pseudo_voigt, two_lorentzian; each entry {largest M, lanes, streamed}, the last",
        "// one open.",
        "#pragma once",
        "",
        "#include <stdint.h>",
        "",
        "namespace bfhip {",
        "",
        "struct TunedRange {",
        "    uint32_t max_points;",
        "    int lanes;",
        "    int streamed;  // 1: points read from memory on every pass",
        "};",
        f"constexpr int kTunedRanges = {width};",
        "constexpr TunedRange kTuned[4][kTunedRanges] = {",
    ]
    for model in MODELS:
        entries = list(rng[model])
        while len(entries) < width:
            entries.append(entries[-1])
        body = ", ".join(f"{{{m if m != 0xFFFFFFFF else '0xffffffffu'}, {g}, {st}}}"
                         for m, g, st in entries)
        lines.append(f"    {{{body}}},  // {model}")
    lines += ["};", "", "}  // namespace bfhip", ""]
    with open(path, "w") as fh:
        fh.write("\n".join(lines))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m batchfit_amd.autotune")
    ap.add_argument("--fits", type=int, default=65536)
    ap.add_argument("--repeat", type=int, default=3)
    ap.add_argument("--write", action="store_true",
                    help="write csrc/bf_tuned.hpp and rebuild the library")
    ap.add_argument("--from-json", help="take the measurements from a saved JSON line "
                                        "instead of measuring")
    args = ap.parse_args(argv)
    if args.from_json:
        saved = json.loads(open(args.from_json).read().strip().splitlines()[-1])
        table = {}
        for key, v in saved["measured"].items():
            model, m = key.split(":")
            table[(model, int(m))] = (str(v["winner"]), dict(v["ms"]))
        args.fits = saved["fits"]
    else:
        table = measure(args.fits, args.repeat)
    rng = ranges(table)
    print(json.dumps({"fits": args.fits,
                      "measured": {f"{k[0]}:{k[1]}": {"winner": v[0], "ms": v[1]}
                                   for k, v in table.items()},
                      "ranges": rng}))
    if args.write:
        write_header(rng, os.path.join(HERE, "csrc", "bf_tuned.hpp"))
        from .build import build
        build()
    return 0


if __name__ == "__main__":
    sys.exit(main())
