"""Builds libmedimage_amd.so (gfx950) in-tree with hipcc.

Every translation unit is compiled with -ffp-contract=off (the tracers'
index/length sequences depend on it) and with packed FP32 instructions
disabled.  Objects are compiled in parallel, then linked.

Then the reference-surface libraries over its C ABI (host code, g++):
libmedimage_goldenc.so (GoldenC's A, AStar, MumfordShah, ...) and
libmedimage_msbeam.so (msbeam's A, AStar), and the two C programs the tests
link against them (tests/c/bin/).
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

SOURCES = ["trace.hip", "cone.hip", "cone_at.hip", "sell.hip", "kernels.hip", "engine.hip", "solver.hip", "ordered.hip", "batch.cpp", "capi.cpp"]

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
    if out == LIB:
        build_compat(verbose)
    return out


COMPAT = {"libmedimage_goldenc.so": "goldenc.cpp", "libmedimage_msbeam.so": "msbeam.cpp"}
CALLERS = {"goldenc_caller": "libmedimage_goldenc.so", "msbeam_caller": "libmedimage_msbeam.so"}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed:\n{r.stderr}")


def build_compat(verbose: bool = False) -> None:
    """The reference-surface shims (found next to libmedimage_amd.so through
    $ORIGIN) and the C callers of the tests (found through a relative rpath,
    so the tree works wherever it is copied)."""
    inc = f"-I{ROOT / 'include'}"
    for so, src in COMPAT.items():
        tmp = PKG / (so + ".tmp")
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", inc, "-o", str(tmp), str(PKG / "compat" / src),
              f"-L{PKG}", "-lmedimage_amd", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, PKG / so)
    bindir = ROOT / "tests" / "c" / "bin"
    bindir.mkdir(parents=True, exist_ok=True)
    rel = os.path.relpath(PKG, bindir)
    for exe, so in CALLERS.items():
        _run(["gcc", "-O2", "-std=c99", "-Wall", inc, "-o", str(bindir / exe), str(ROOT / "tests" / "c" / (exe + ".c")),
              f"-L{PKG}", f"-l:{so}", f"-Wl,-rpath,$ORIGIN/{rel}"])
    if verbose:
        print("built " + ", ".join(list(COMPAT) + [f"tests/c/bin/{e}" for e in CALLERS]))


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
