"""The C ABI library: built for gfx950, loads, exports every entry point
include/medimage_amd.h declares, and refuses to run without a device (no
CPU fallback).  No GPU needed."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared(header):
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(mi_\w+)\s*\(", text, flags=re.M)))


def test_headers_exist():
    assert HEADERS, "include/*.h missing"


def test_library_exports_every_declared_symbol():
    import medimage_amd as mi
    lib = C.CDLL(str(mi.LIB_PATH))
    names = [n for h in HEADERS if h.name == "medimage_amd.h" for n in declared(h)]
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("lib,header", [("libmedimage_goldenc.so", "medimage_goldenc.h"),
                                        ("libmedimage_msbeam.so", "medimage_msbeam.h")])
def test_compat_libraries_export_the_reference_surface(lib, header):
    """The GoldenC / msbeam compatibility libraries export every function
    and global their headers declare, under the reference's own names."""
    import medimage_amd as mi
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / header).read_text(), flags=re.S)
    funcs = re.findall(r"^void\s+(\w+)\s*\(", text, flags=re.M)
    data = [n for decl in re.findall(r"^extern\s+double\s+([^;]+);", text, flags=re.M)
            for n in re.findall(r"(\w+)\s*\[", decl)]
    assert funcs
    out = subprocess.run(["nm", "-D", "--defined-only", str(mi.LIB_PATH.parent / lib)], capture_output=True,
                         text=True).stdout
    exported = {line.split()[-1]: line.split()[-2] for line in out.splitlines() if len(line.split()) >= 3}
    for n in funcs:
        assert exported.get(n) == "T", (n, exported.get(n))
    for n in data:
        assert exported.get(n) in ("B", "D"), (n, exported.get(n))


def test_library_is_gfx950_code_object():
    import medimage_amd as mi
    out = subprocess.run(["/opt/rocm/lib/llvm/bin/llvm-readelf", "-S", str(mi.LIB_PATH)], capture_output=True,
                         text=True).stdout
    assert ".hip_fatbin" in out
    blob = mi.LIB_PATH.read_bytes()
    assert b"gfx950" in blob


def test_no_device_means_error_not_fallback():
    import medimage_amd as mi
    lib = mi.lib()
    h = C.c_void_p()
    rc = lib.mi_context_create(0, C.byref(h))
    if rc == mi.MI_OK:
        lib.mi_context_destroy(h)
        pytest.skip("a device is present")
    assert rc == mi.MI_E_DEVICE
    assert lib.mi_last_error()


def test_null_arguments_rejected():
    import medimage_amd as mi
    lib = mi.lib()
    assert lib.mi_context_create(0, None) == mi.MI_E_ARG
    assert lib.mi_forward(None, None, None, None) in (mi.MI_E_ARG,)
