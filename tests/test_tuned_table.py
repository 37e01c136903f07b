"""The measured mapping table (csrc/bf_tuned.hpp) and the autotuner's host
logic: ranges are increasing, end open with whole waves, and never ask a
partial wave for more than its largest register tile (32 points per lane)."""
import os
import re

from batchfit_amd import autotune

HEADER = os.path.join(os.path.dirname(autotune.__file__), "csrc", "bf_tuned.hpp")


def parse_table():
    text = open(HEADER).read()
    rows = re.findall(r"\{(\{.*\})\},\s*//\s*(\w+)", text)
    out = {}
    for body, model in rows:
        entries = re.findall(r"\{(0x[0-9a-f]+u|\d+),\s*(\d+),\s*([01])\}", body)
        out[model] = [(int(m.rstrip("u"), 0), int(g), int(st)) for m, g, st in entries]
    return out


def test_table_covers_every_model_and_m():
    table = parse_table()
    assert list(table) == list(autotune.MODELS)
    for model, entries in table.items():
        assert entries[-1][0] == 0xFFFFFFFF, model
        bounds = [m for m, _, _ in entries]
        assert bounds == sorted(bounds), model
        for m, g, streamed in entries:
            assert g in (4, 8, 16, 32, 64), model
            if g == 4:
                assert streamed, (model, m)  # 4-lane groups only stream
            if g < 64 and not streamed:
                assert m <= 32 * g, (model, m, g)


def lookup(entries, m):
    for bound, g, streamed in entries:
        if m <= bound:
            return autotune.label(g, streamed)
    return "64"


def test_ranges_merge_and_clamp():
    # Synthetic measurements: 16 lanes win to 512, 32 to 1024, then 64.
    winners = {m: "16" if m <= 512 else "32" if m <= 1024 else "64" for m in autotune.LADDER}
    table = {(model, m): (g, {g: 1.0}) for model in autotune.MODELS for m, g in winners.items()}
    rng = autotune.ranges(table)
    for model in autotune.MODELS:
        entries = rng[model]
        for m in autotune.LADDER:
            assert lookup(entries, m) == winners[m], (model, m)
        # The 16-lane range ends where its largest tile does, not at the midpoint 640.
        assert entries[0] == (512, 16, 0)
        assert entries[-1] == (0xFFFFFFFF, 64, 0)


def test_ranges_streamed_winners_are_open():
    # Streamed 16-lane groups win from 1536 on: no tile cap, and the open
    # range is theirs.
    winners = {m: "16" if m <= 512 else "64" if m <= 1024 else "16s" for m in autotune.LADDER}
    table = {(model, m): (g, {g: 1.0}) for model in autotune.MODELS for m, g in winners.items()}
    rng = autotune.ranges(table)
    for model in autotune.MODELS:
        assert rng[model][-1] == (0xFFFFFFFF, 16, 1)
        for m in autotune.LADDER:
            assert lookup(rng[model], m) == winners[m], (model, m)


def test_ranges_append_open_whole_wave_range():
    table = {(model, m): ("16" if m <= 512 else "32", {}) for model in autotune.MODELS
             for m in autotune.LADDER if m <= 1024}
    for model in autotune.MODELS:
        for m in autotune.LADDER:
            table.setdefault((model, m), ("32", {}))
    # 32-lane tiles "win" beyond their capacity: the table must still end open.
    rng = autotune.ranges(table)
    for model in autotune.MODELS:
        assert rng[model][-1] == (0xFFFFFFFF, 64, 0)
        assert all(m <= 32 * g for m, g, st in rng[model] if g < 64 and not st)


def test_header_round_trip(tmp_path):
    winners = {m: "16" if m <= 512 else "32s" for m in autotune.LADDER}
    table = {(model, m): (g, {g: 1.0}) for model in autotune.MODELS for m, g in winners.items()}
    rng = autotune.ranges(table)
    path = tmp_path / "bf_tuned.hpp"
    autotune.write_header(rng, str(path))
    text = path.read_text()
    assert "{512, 16, 0}, {0xffffffffu, 32, 1}" in text


def test_ranges_noise_does_not_split_the_table():
    # Alternating winners 2 % apart cost less than the switches: one mapping.
    table = {}
    for model in autotune.MODELS:
        for i, m in enumerate(autotune.LADDER):
            a, b = (1.0, 1.02) if i % 2 else (1.02, 1.0)
            table[(model, m)] = ("16" if a < b else "16s", {"16": a, "16s": b})
    rng = autotune.ranges(table)
    for model in autotune.MODELS:
        assert rng[model] == [(0xFFFFFFFF, 16, 1)], rng[model]
