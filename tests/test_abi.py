"""CPU-side checks of the C ABI: the library builds, loads and exports every
symbol include/gcdt.h declares (no compute calls without a GPU)."""
import os
import re

import importlib

L = importlib.import_module("gcdt_amd.lib")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "gcdt.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(gcdt_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ["gcdt_solve", "gcdt_get_sample_points", "gcdt_raw_mesh_before", "gcdt_raw_mesh_after",
              "gcdt_get_bitmap", "gcdt_add_constraints", "gcdt_try_fast_delaunay", "gcdt_delaunay_new"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert set(_declared()) <= set(L.SIGNATURES)
