"""Building and locating the native libraries (in-tree, no JIT cache).

``libfc_chol.so``     the C-ABI supernodal Cholesky: host symbolic phase + HIP numeric phase
                      and solves for gfx950 (csrc/*.cpp, csrc/*.hip)
``libfc_problems.so`` host-only synthetic inputs: meshes, element matrices, assembly, ordering

Both are compiled by ``build_all()`` (``__graft_entry__.build()`` calls it). Loading never
builds: a missing library is an error, so the GPU path never runs on anything but the built
HIP code.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
CHOL = os.path.join(HERE, "libfc_chol.so")
PROBLEMS = os.path.join(HERE, "libfc_problems.so")


def build_all(jobs: int = 8) -> None:
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", CSRC], check=True)


def problems_library() -> str:
    if not os.path.exists(PROBLEMS):
        raise RuntimeError(f"{PROBLEMS} is not built: run __graft_entry__.build() (make -C {CSRC})")
    return PROBLEMS


def chol_library() -> str:
    if not os.path.exists(CHOL):
        raise RuntimeError(f"{CHOL} is not built: run __graft_entry__.build() (make -C {CSRC})")
    return CHOL
