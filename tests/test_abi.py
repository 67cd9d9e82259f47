"""CPU checks of the C ABI boundary: the library loads, exports every symbol that
include/dpmd/dpmd.h declares, and the Python binding covers them.  No compute call needs a GPU
here; without one, context creation must fail with a HIP status and a message."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from dpmd_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols(header="dpmd.h"):
    txt = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "dpmd" / header).read_text(), flags=re.S)
    return set(re.findall(r"\b(dpmd_[a-z0-9_]+)\s*\(", txt))


def test_every_header_is_checked():
    assert {h.name for h in (ROOT / "include").rglob("*.h")} == {"dpmd.h", "dpmd_h5.h"}


def test_header_declares_entry_points():
    names = declared_symbols()
    assert len(names) > 50
    for must in ("dpmd_ctx_create", "dpmd_compute_forces", "dpmd_fire_minimize", "dpmd_vv_run",
                 "dpmd_init_neighbors", "dpmd_get_neighbors_csr", "dpmd_compute_pressure"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (dpmd_\w+)", out))
    assert declared_symbols() <= exported


def test_python_binding_covers_declarations():
    assert declared_symbols() <= set(L.PROTOS)


def test_version_and_error_path_without_gpu():
    lib = L.lib()
    assert lib.dpmd_version().startswith(b"dpmd_amd")
    p = C.c_void_p()
    rc = lib.dpmd_ctx_create(C.byref(p), 0, None)
    if rc == 0:
        lib.dpmd_ctx_destroy(p)
        pytest.skip("a GPU is present")
    assert rc == 3
    assert lib.dpmd_last_error(None)
    # bad kind is an argument error, reported before any device call
    assert lib.dpmd_ctx_create(C.byref(p), 7, None) == 1
    assert b"kind" in lib.dpmd_last_error(None)


@pytest.mark.skipif(not (ROOT / "dpmd_amd" / "libdpmd_h5.so").exists(), reason="libdpmd_h5 not built (no HDF5)")
def test_h5_library_exports_and_binding():
    from dpmd_amd import h5
    names = declared_symbols("dpmd_h5.h")
    assert "dpmd_h5_load" in names and len(names) >= 15
    lib = h5.lib()
    assert all(hasattr(lib, n) for n in names)
    out = subprocess.run(["nm", "-D", "--defined-only", str(h5.LIB_PATH)], capture_output=True, text=True).stdout
    assert names <= set(re.findall(r"\bT (dpmd_h5_\w+)", out))
    assert names <= set(h5._PROTOS)
