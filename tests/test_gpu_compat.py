"""The reference-surface libraries on the GPU, driven by C programs that
link them the way the reference's own drivers link proj.c:

* tests/c/goldenc_caller.c (GoldenC's calling pattern) against
  libmedimage_goldenc.so: A(gDelta, init_f) and the whole
  MumfordShah(1000, 0.002, 1e-4), 10 x (10 CG + 10 SD), equal bit for bit to
  the reference program's own run (tests/golden/goldenc_ms.npz);
* tests/c/msbeam_caller.c against libmedimage_msbeam.so: msbeam's A and
  AStar equal bit for bit to the oracle's restatement (float sums in
  msbeam's loop order; pinned to the compiled reference in
  tests/test_oracle.py).
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent / "c" / "bin"


def run(exe, *args):
    print(f"[compat] {exe.name} {' '.join(map(str, args))}", flush=True)
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr


def test_goldenc_caller_end_to_end(tmp_path, golden):
    f0 = golden("blockrecon_std.npz")["f"].astype(np.float64)
    z = golden("goldenc_ms.npz")
    ph = tmp_path / "phantom.txt"
    np.savetxt(ph, f0, fmt="%.17g")
    out = tmp_path / "out.bin"
    run(BIN / "goldenc_caller", ph, out)
    o = np.fromfile(out, np.float64)
    nr, npx = 180 * 767, 512 * 512
    g, f, v = o[:nr].reshape(180, 767), o[nr:nr + npx].reshape(512, 512), o[nr + npx:].reshape(512, 512)
    assert np.array_equal(g.view(np.uint64), z["g"].view(np.uint64))
    assert np.array_equal(f.view(np.uint64), z["f"].view(np.uint64))
    assert np.array_equal(v.view(np.uint64), z["v"].view(np.uint64))


def test_msbeam_caller_projector_pair(tmp_path, golden):
    f = golden("blockrecon_std.npz")["f"].astype(np.float32)
    rs = np.random.RandomState(11)
    f = (f + rs.rand(*f.shape).astype(np.float32) * np.float32(0.25)).astype(np.float32)
    src = tmp_path / "f.bin"
    f.tofile(src)
    out = tmp_path / "out.bin"
    run(BIN / "msbeam_caller", src, out)
    o = np.fromfile(out, np.float32)
    g, b = o[:180 * 768].reshape(180, 768), o[180 * 768:].reshape(512, 512)
    og = O.par_fp_f32(f, 180, 768, 0)
    ob = O.par_bp_f32(og, 512, 0)
    assert np.array_equal(g.view(np.uint32), og.view(np.uint32))
    assert np.array_equal(b.view(np.uint32), ob.view(np.uint32))
