"""CPU checks of the NPY boundary (bf_npy_inspect, bf_npy_chunk_fits): the
header rules of the reference's reader (src/cpu/npy.cpp:51-84: v1.0, <f4/<f8/
<u4, C order, rank 1/2, keys in order, payload length) on files NumPy writes
and on hand-made broken ones, and the chunk arithmetic."""
import os

import numpy as np
import pytest

from batchfit_amd import npy


def header_bytes(text: str, magic=b"\x93NUMPY\x01\x00"):
    body = text.encode()
    pad = (16 - (10 + len(body) + 1) % 16) % 16
    body = body + b" " * pad + b"\n"
    return magic + len(body).to_bytes(2, "little") + body


@pytest.mark.parametrize("dtype,shape", [("<f4", (7,)), ("<f8", (3, 5)), ("<u4", (4,)),
                                         ("<f4", (0, 9)), ("<f8", (2, 0))])
def test_inspect_numpy_files(tmp_path, dtype, shape):
    path = tmp_path / "a.npy"
    np.save(path, np.zeros(shape, dtype))
    h = npy.inspect(path)
    assert h.shape == shape and h.dtype == np.dtype(dtype)
    assert h.offset + int(np.prod(shape)) * np.dtype(dtype).itemsize == os.path.getsize(path)


@pytest.mark.parametrize("text,payload,message", [
    ("{'descr': '>f4', 'fortran_order': False, 'shape': (2,), }", 8, "unsupported header"),
    ("{'descr': '<f4', 'fortran_order': True, 'shape': (2,), }", 8, "unsupported header"),
    ("{'descr': '<f4', 'fortran_order': False, 'shape': (2, 2, 2), }", 32, "unsupported header"),
    ("{'fortran_order': False, 'descr': '<f4', 'shape': (2,), }", 8, "unsupported header"),
    ("{'descr': '<i8', 'fortran_order': False, 'shape': (2,), }", 16, "unsupported header"),
    ("{'descr': '<f4', 'fortran_order': False, 'shape': (3,), }", 8, "payload length"),
])
def test_inspect_rejects(tmp_path, text, payload, message):
    path = tmp_path / "bad.npy"
    path.write_bytes(header_bytes(text) + b"\0" * payload)
    with pytest.raises(ValueError, match=message):
        npy.inspect(path)


def test_inspect_rejects_magic_and_missing(tmp_path):
    path = tmp_path / "v2.npy"
    path.write_bytes(header_bytes("{'descr': '<f4', 'fortran_order': False, 'shape': (1,), }",
                                  magic=b"\x93NUMPY\x02\x00") + b"\0" * 4)
    with pytest.raises(ValueError, match="magic"):
        npy.inspect(path)
    with pytest.raises(ValueError, match="open"):
        npy.inspect(tmp_path / "missing.npy")


def test_chunk_fits():
    # Explicit chunks are clamped to N; budgets round down to multiples of 256.
    assert npy.chunk_fits(1000, 129, 3, chunk=300) == 300
    assert npy.chunk_fits(100, 129, 3, chunk=300) == 100
    c = npy.chunk_fits(10 ** 7, 1024, 4, 4, budget_bytes=1 << 30)
    per_fit = 3 * (1024 * 8 + 4 * 4 + 32 + 12)
    assert c == (1 << 30) // per_fit // 256 * 256
    assert npy.chunk_fits(10 ** 7, 1024, 4, 8, in_flight=1, uncertainties=True) > c
    assert npy.chunk_fits(0, 1024, 4) == 0
    assert npy.chunk_fits(10, 1024, 9) == 0
