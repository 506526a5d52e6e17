"""CPU calibration of the reference programs against the repository's CPU
restatement (run in a container where /root/reference and `make -C oracle
ref ref-native` exist; nothing here runs on the GPU host).

Times, on this host:
* GoldenC (origin/GoldenC/src): one alternation of MumfordShah (10 CG
  image steps + 10 SD edge steps on 512 x 512, 180 views x 767 bins, the
  reference's config.h), through its own minimize_in_IMAGE_variable /
  minimize_in_EDGE_variable, built with strict IEEE flags and with its
  Makefile's native flags; against the oracle's restatement (or_gc_*).
* msbeam/CPU_OpenMP: its main() (phantom, A, then MSbeam(): 10 per-ray AT
  sweeps at 512 x 512, 180 views x 768 rays; main prints the sweep time,
  msbeam/CPU_OpenMP/src/main.c:196-202) at 1 and 8 threads, both builds;
  against the oracle's one-thread restatement (or_msbeam_solve).

Prints one JSON object with the timings.
"""
import ctypes as C
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
sys.path.insert(0, str(ROOT / "tests"))
import oracle_py as O  # noqa: E402

N, NV, NB = 512, 180, 767


def goldenc_ref(lib_name):
    lib = C.CDLL(str(HERE / "_ref" / lib_name))
    f0 = np.load(ROOT / "tests" / "golden" / "blockrecon_std.npz")["f"].astype(np.float64)
    init_f = (C.c_double * (N * N)).in_dll(lib, "init_f")
    C.memmove(init_f, f0.ctypes.data, f0.nbytes)
    gd = (C.c_double * (NV * NB)).in_dll(lib, "gDelta")
    fa = (C.c_double * (N * N)).in_dll(lib, "f")
    va = (C.c_double * (N * N)).in_dll(lib, "v")
    lib.A(gd, init_f)
    C.memset(fa, 0, N * N * 8)
    np.frombuffer(va, np.float64)[:] = 1.0
    lib.minimize_in_IMAGE_variable.argtypes = [C.c_void_p] * 4 + [C.c_double, C.c_double, C.c_int]
    lib.minimize_in_EDGE_variable.argtypes = [C.c_void_p] * 3 + [C.c_double] * 3 + [C.c_int]
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)
    try:
        t0 = time.perf_counter()
        lib.minimize_in_IMAGE_variable(fa, gd, fa, va, 1000.0, 1e-4, 10)
        lib.minimize_in_EDGE_variable(va, fa, va, 1000.0, 0.002, 1e-4, 10)
        dt = time.perf_counter() - t0
    finally:
        os.dup2(saved, 1)
        os.close(devnull)
    return dt


def goldenc_port():
    f0 = np.load(ROOT / "tests" / "golden" / "blockrecon_std.npz")["f"].astype(np.float64)
    g = O.gc_fp(f0)
    t0 = time.perf_counter()
    O.mumford_shah(O.op_goldenc(), g, 1000.0, 0.002, 1e-4, 1, 10, 10)
    return time.perf_counter() - t0


def msbeam_ref(lib_name, threads):
    """main() in a child process (it writes its files into the cwd)."""
    code = (f"import ctypes as C; lib = C.CDLL({str(HERE / '_ref' / lib_name)!r}); "
            f"argv = (C.c_char_p * 2)(b'msbeam', b'{threads}'); lib.main(2, argv)")
    with tempfile.TemporaryDirectory() as d:
        out = subprocess.run([sys.executable, "-c", code], cwd=d, capture_output=True, text=True, check=True).stdout
    return float(re.search(r"TIME USED: ([0-9.]+)", out).group(1))


def msbeam_port():
    """The restatement's 10 sweeps from the sinogram of the reference's own
    main() (its phantom, g = A f), kept in the msbeam fixture."""
    z = np.load(ROOT / "tests" / "golden" / "msbeam_ref.npz")
    t0 = time.perf_counter()
    O.msbeam_solve(z["g"], 512, 10.0, 0.35, 1e-3, 10)
    return time.perf_counter() - t0


def main():
    out = {"host": {"cpus": os.cpu_count(), "model": next((l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                                                            if l.startswith("model name")), "")}}
    out["goldenc_one_alternation_s"] = {
        "reference_strict": goldenc_ref("libref_goldnc.so"),
        "reference_native": goldenc_ref("libref_goldnc_native.so"),
        "port": goldenc_port(),
    }
    out["msbeam_10_sweeps_s"] = {
        "reference_strict_1t": msbeam_ref("libref_msbeam.so", 1),
        "reference_strict_8t": msbeam_ref("libref_msbeam.so", 8),
        "reference_native_1t": msbeam_ref("libref_msbeam_native.so", 1),
        "reference_native_8t": msbeam_ref("libref_msbeam_native.so", 8),
        "port_1t": msbeam_port(),
    }
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
