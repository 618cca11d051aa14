"""Build libtfm_hip.so for gfx950 in-tree with hipcc.

The library is placed next to the Python package (torchfem/libtfm_hip.so) so
that it is loaded from the source tree.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
OUT = HERE / "torchfem" / "libtfm_hip.so"
BUILD_DIR = HERE / "_build"
SOURCES = ["assembly.hip", "solver.hip", "capi.hip", "loads.hip", "krylov.hip", "pattern.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # keep packed FP32 instructions out of the generated code
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed: {' '.join(cmd)}")
    if verbose and res.stderr.strip():
        # hipcc's host pass warns about the unknown packed-fp32 feature
        lines = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if lines:
            print("\n".join(lines), flush=True)


def build(verbose: bool = True) -> Path:
    """Compile every HIP source for gfx950 and link libtfm_hip.so."""
    BUILD_DIR.mkdir(exist_ok=True)
    objs = []

    def compile_one(src: str) -> Path:
        obj = BUILD_DIR / (Path(src).stem + ".o")
        _run([HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)], verbose)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    _run(
        [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)],
        verbose,
    )
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
