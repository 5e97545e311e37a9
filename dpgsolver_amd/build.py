"""Build libdgx.so (hipcc, gfx950) in-tree."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build_lib(force: bool = True, jobs: int = 8) -> str:
    csrc = os.path.join(HERE, "csrc")
    cmd = ["make", "-s", "-C", csrc, f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return os.path.join(HERE, "libdgx.so")


if __name__ == "__main__":
    print(build_lib(force="--force" in sys.argv))
