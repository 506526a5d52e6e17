"""Writes the golden fixtures under tests/golden/ (run in the build container,
where /root/reference and the reference libraries built by `make -C oracle ref`
exist).

* blockrecon_std.npz: blockrecon/phantom/std_f.dat (the normalised 512 x 512
  phantom) and std_g.dat (A(std_f) printed by the reference to 6 significant
  digits, blockrecon/src/blockrecon.cpp:374-379), parsed to arrays.
* goldenc_ms.npz: the reference GoldenC program (origin/GoldenC/src, strict
  IEEE build in oracle/_ref) run end to end: gDelta = A(std_f), then
  MumfordShah(1000, 0.002, 1e-4) with 10 x (10 CG + 10 SD) iterations
  (main.c:60-61, config.h:6-17).  Stores gDelta, f and v.
* msbeam_ref.npz: the reference msbeam OpenMP program at one thread
  (main(argc=2, "1")): its own phantom, g = A(f), 10 AT sweeps; stores g, f, v.
"""
import ctypes as C
import os
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REFSRC = Path("/root/reference")
REFLIB = ROOT / "oracle" / "_ref"


def blockrecon_std():
    f = np.loadtxt(REFSRC / "blockrecon/phantom/std_f.dat", dtype=np.float64)
    g = np.array(open(REFSRC / "blockrecon/phantom/std_g.dat").read().split(), dtype=np.float64).reshape(180, 768)
    np.savez_compressed(HERE / "blockrecon_std.npz", f=f.astype(np.float32), g=g)


def goldenc_ms():
    lib = C.CDLL(str(REFLIB / "libref_goldnc.so"))
    f0 = np.loadtxt(REFSRC / "blockrecon/phantom/std_f.dat", dtype=np.float64)
    init_f = (C.c_double * (512 * 512)).in_dll(lib, "init_f")
    C.memmove(init_f, f0.ctypes.data, f0.nbytes)
    gd = (C.c_double * (180 * 767)).in_dll(lib, "gDelta")
    lib.A(gd, init_f)
    lib.MumfordShah.argtypes = [C.c_double] * 3
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)
    try:
        lib.MumfordShah(1000.0, 0.002, 0.0001)
    finally:
        os.dup2(saved, 1)
    f = np.frombuffer((C.c_double * (512 * 512)).in_dll(lib, "f"), np.float64).reshape(512, 512).copy()
    v = np.frombuffer((C.c_double * (512 * 512)).in_dll(lib, "v"), np.float64).reshape(512, 512).copy()
    g = np.frombuffer(gd, np.float64).reshape(180, 767).copy()
    np.savez_compressed(HERE / "goldenc_ms.npz", g=g, f=f, v=v)


def msbeam_ref():
    lib = C.CDLL(str(REFLIB / "libref_msbeam.so"))
    argv = (C.c_char_p * 2)(b"msbeam", b"1")
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        os.chdir(d)
        devnull = os.open(os.devnull, os.O_WRONLY)
        saved = os.dup(1)
        os.dup2(devnull, 1)
        try:
            lib.main(2, argv)
        finally:
            os.dup2(saved, 1)
            os.chdir(cwd)
    n = 512
    g = np.frombuffer((C.c_float * (180 * 768)).in_dll(lib, "g"), np.float32).reshape(180, 768).copy()
    f = np.frombuffer((C.c_float * (n * n)).in_dll(lib, "f"), np.float32).reshape(n, n).copy()
    v = np.frombuffer((C.c_float * (n * n)).in_dll(lib, "v"), np.float32).reshape(n, n).copy()
    np.savez_compressed(HERE / "msbeam_ref.npz", g=g, f=f, v=v)


if __name__ == "__main__":
    which = sys.argv[1:] or ["blockrecon", "goldenc", "msbeam"]
    if "blockrecon" in which:
        blockrecon_std()
    if "msbeam" in which:
        msbeam_ref()
    if "goldenc" in which:
        goldenc_ms()
