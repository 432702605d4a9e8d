"""The C-ABI library loads and exports every entry point include/zr_mpm.h
declares (CPU only: no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zr_mpm.h")


def declared_functions(path=HEADER):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "zr_create" in names and "zr_advance_one_time_step" in names


def test_library_exports_every_declared_symbol():
    from ziran2019_amd import engine
    if not os.path.exists(engine.LIB_PATH):
        pytest.fail("libzr_mpm.so is not built; run python -m ziran2019_amd.build")
    out = subprocess.run(["nm", "-D", "--defined-only", engine.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    assert sorted(engine.EXPORTS) == declared_functions()


def test_library_loads_without_gpu():
    from ziran2019_amd import engine
    lib = engine.load_library()
    for f in declared_functions():
        assert isinstance(getattr(lib, f), ctypes._CFuncPtr)
