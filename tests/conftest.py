"""Shared fixtures. Tests marked ``gpu`` need an MI355X; the rest run on CPU.

The CPU oracle (oracle/liboracle.so) is the checker for every parity test; it
is built here on first use if missing (a few seconds of g++). The HIP library
must already be built (``python -m batchfit_amd.build`` or
``__graft_entry__.build()``); GPU tests fail, not skip, when it is missing.
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for path in (ROOT, GOLDEN):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def oracle_port():
    from oracle import oracle
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"),
                        os.path.join(ROOT, "oracle", "liboracle.so")], check=True)
    return oracle.load("port")


@pytest.fixture(scope="session")
def golden():
    return {name: np.load(os.path.join(GOLDEN, f"golden_{name}.npz"))
            for name in ("config1", "sets", "failures")}


@pytest.fixture(scope="session")
def hip():
    """The HIP package with the library loaded and a gfx950 device present."""
    import batchfit_amd
    batchfit_amd._native.library()  # raises if the library is not built
    assert batchfit_amd.hip_available(), "no usable gfx950 device"
    return batchfit_amd
