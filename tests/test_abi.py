"""CPU-side checks of the C ABI: the library is built for gfx950, loads, and
exports every entry point include/polymer_gpu.h declares.  No compute call is
made (there is no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "polymer_gpu.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^int\s+(pg_\w+)\s*\(", text, re.M)))


def test_header_and_python_list_agree():
    import polymer_amd
    assert declared_symbols() == sorted(polymer_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    import polymer_amd
    if not os.path.exists(polymer_amd.LIB_PATH):
        pytest.fail("libpolymer_gpu.so not built")
    L = ctypes.CDLL(polymer_amd.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_library_has_gfx950_code():
    import polymer_amd
    out = subprocess.run(["strings", polymer_amd.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_last_error_without_init():
    import polymer_amd
    L = polymer_amd.load()
    buf = ctypes.create_string_buffer(64)
    assert L.pg_last_error(buf, 64) == 0
    # a compute call before pg_init is refused, not run on a CPU path
    h = ctypes.c_void_p()
    rc = L.pg_graph_rmat(4, 10, 0.57, 0.19, 0.19, 1, 0, 0, ctypes.byref(h))
    assert rc == -4  # PG_ERR_STATE
    L.pg_last_error(buf, 64)
    assert b"pg_init" in buf.value
