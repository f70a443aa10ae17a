"""The C-ABI library loads and exports every entry point declared in
include/*.h (no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

import dust2_amd
from dust2_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|const char\*)\s+(dust2_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared()
    assert "dust2_filter_sir_alloc" in names
    assert "dust2_filter_run" in names
    assert len(names) > 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert missing == []


def test_python_mirror_binds_declared_symbols():
    assert set(_lib.declared_symbols()) <= declared()


def test_version():
    assert dust2_amd.lib().dust2_amd_version() == b"0.1.0"
