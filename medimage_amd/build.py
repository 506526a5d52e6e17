"""Builds libmedimage_amd.so (gfx950) in-tree with hipcc.

Every translation unit is compiled with -ffp-contract=off (the tracers'
index/length sequences depend on it) and with packed FP32 instructions
disabled.  Objects are compiled in parallel, then linked.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_obj"
LIB = PKG / "libmedimage_amd.so"

SOURCES = ["trace.hip", "cone.hip", "cone_at.hip", "sell.hip", "kernels.hip", "engine.hip", "solver.hip", "batch.cpp", "capi.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    # per-view tables call libm sin and cos separately (no sincos merge)
    "-fno-builtin-sin", "-fno-builtin-cos", "-fno-builtin-sinf", "-fno-builtin-cosf",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
]


def _compile(src: str, obj: Path = OBJ, defines=()) -> Path:
    out = obj / (Path(src).stem + ".o")
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-c", str(CSRC / src), "-o", str(out)]
    if src.endswith(".cpp"):
        cmd.insert(1, "-xhip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


def build(verbose: bool = False, out: Path = LIB, obj: Path = OBJ, defines=()) -> Path:
    """Compiles every source and links `out`; `defines` (NAME=VALUE) select
    compile-time kernel variants."""
    obj.mkdir(parents=True, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, obj, defines), SOURCES))
    tmp = out.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
