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


def test_cpp_dropin_header_compiles_against_reference_call_sites(tmp_path):
    """The reference's host call sequence (kernel.cu:5399-5446, addConstraints.h:5-6)
    type-checks against include/gcdt.hpp unchanged."""
    import shutil
    import subprocess

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    src = tmp_path / "main.cpp"
    src.write_text(r'''
#include "gcdt.hpp"
void run(void* tmp, size_t bytes, Point2d* points, int numPoints, int2* cons, int numCons) {
    Point2d* samplePoints; int* sampleId; int numSample, numTriangles; int3 *tris, *adjTris, *newTris, *sampleTris;
    int* sons; int *bitmapOffsets, *bitmapTris; int3* consTris; int numConsTris = 0;
    double t = getSamplePoints(tmp, bytes, points, numPoints, samplePoints, sampleId, numSample);
    t += rawMeshBefore(tmp, bytes, samplePoints, tris, sons, numTriangles, numSample);
    t += filterTriangles(tmp, bytes, tris, sampleTris, nullptr, nullptr, sons, numTriangles);
    t += delaunaySample(tmp, bytes, samplePoints, sampleTris, adjTris, numSample + 4, numTriangles);
    t += getBitmap(tmp, bytes, samplePoints, sampleTris, bitmapOffsets, bitmapTris, numTriangles);
    convertSample(sampleTris, sampleId, numSample, numPoints, numTriangles);
    t += rawMeshAfter(tmp, bytes, points, tris, adjTris, sons, bitmapOffsets, bitmapTris, numTriangles, numPoints);
    t += addConstraints(tmp, bytes, points, tris, adjTris, sons, bitmapOffsets, bitmapTris, cons, consTris, numCons,
                        numPoints, numTriangles, numConsTris);
    t += addConstraints(tmp, bytes, points, tris, adjTris, sons, cons, numCons, numPoints, numTriangles);
    t += delaunayNew(tmp, bytes, points, newTris, cons, numPoints + 4, numTriangles, numCons);
    t += tryFastDelaunay(tmp, bytes, points, newTris, adjTris, numPoints + 4, numTriangles);
    rawMesh(tmp, bytes, points, tris, adjTris, sons, numTriangles, numPoints);
    (void)t;
}
''')
    r = subprocess.run([hipcc, "-fsyntax-only", "-std=c++17", "--offload-arch=gfx950", "-x", "hip",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
