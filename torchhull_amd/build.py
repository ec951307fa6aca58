"""In-tree build of the HIP library (hipcc, gfx950) via torchhull_amd/csrc/Makefile."""

from __future__ import annotations

import os
import pathlib
import subprocess

CSRC = pathlib.Path(__file__).resolve().parent / "csrc"


def build(jobs: int = 4, force: bool = True) -> pathlib.Path:
    """Compile every HIP source for gfx950 and link libtorchhull_hip.so; returns its path."""
    env = dict(os.environ)
    args = ["make", "-C", str(CSRC), f"-j{jobs}"]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, env=env)
    return CSRC.parent / "_build" / "libtorchhull_hip.so"
