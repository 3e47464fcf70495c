"""The C ABI library loads and exports every entry point that
include/neorx.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import neorx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "neorx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(neorx_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = neorx.load()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", neorx.LIBPATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (neorx_\w+)", out))
    assert set(names) <= exported, set(names) - exported


def test_library_carries_gfx950_code_object():
    sec = subprocess.run(["readelf", "-S", neorx.LIBPATH], capture_output=True, text=True).stdout
    assert ".hip_fatbin" in sec
    blob = open(neorx.LIBPATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_no_cpu_fallback_without_device():
    """Without a usable gfx950 device every entry point reports an error."""
    lib = neorx.load()
    rc = lib.neorx_init(0)
    if rc != 0:
        assert rc == -1
        assert b"device" in lib.neorx_last_error()
        h = ctypes.c_int64()
        z = (ctypes.c_double * 8)()
        i = (ctypes.c_int64 * 8)(1, 1, 1, 1, 1, 1, 1, 1)
        rc2 = lib.neorx_system_create(1, 1, 1, 1, 1, ctypes.c_double(1.0), z, z, z, z, i, i, i, z, z, z, z, i, i, z,
                                      z, ctypes.byref(h))
        assert rc2 != 0
