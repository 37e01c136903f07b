"""The numpy Philox4x32-10 that checks the device generator reproduces the
published known-answer vectors of Random123 (kat_vectors, philox4x32 with 10
rounds), so a wrong multiplier, key schedule or word order fails here."""
import numpy as np

from philox import philox4x32_10, unit


def test_philox_known_answers():
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF),
         (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        got = philox4x32_10(np.array([ctr], np.uint32), key)[0]
        assert [int(v) for v in got] == list(want)


def test_unit_interval():
    w = np.array([0, 0xFF, 0x100, 0xFFFFFFFF], np.uint32)
    u = unit(w)
    assert u[0] == u[1] == np.float32(0.5 * 2.0 ** -24)
    assert u[2] == np.float32(1.5 * 2.0 ** -24)
    # float32 rounds 2^24 - 0.5 up: the top of the range is exactly 1.
    assert u[3] == np.float32(1.0)
