"""In-tree build of libuwshcu_hip.so (hipcc, gfx950) via csrc/Makefile."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")


def build(arch: str = "gfx950", jobs: int = 8, force: bool = False) -> str:
    """Compile every HIP source for ``arch``; returns the path of the shared library."""
    env = dict(os.environ)
    if force:
        subprocess.run(["make", "-s", "-C", CSRC, "clean"], check=True, env=env)
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}", f"ARCH={arch}"], check=True, env=env)
    return os.path.join(HERE, "libuwshcu_hip.so")
