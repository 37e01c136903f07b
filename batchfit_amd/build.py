"""Build libbatchfit_hip.so in-tree for gfx950: ``python -m batchfit_amd.build``."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 5, arch: str = "gfx950") -> str:
    """Compile every translation unit under csrc/ with hipcc and link the library."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", os.path.join(HERE, "csrc"), f"ARCH={arch}"],
                   check=True)
    return os.path.join(HERE, "libbatchfit_hip.so")


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else 5))
