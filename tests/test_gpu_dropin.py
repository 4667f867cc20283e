"""GPU parity of the reference's host pipeline run through the drop-in entry
points, in the order of the reference's main (kernel.cu:5399-5446):
getSamplePoints -> rawMeshBefore -> filterTriangles -> delaunaySample ->
getBitmap -> convertSample -> rawMeshAfter -> addConstraints ->
filterTriangles (+ append) -> delaunayNew.  The final triangle list must be the
oracle's CDT (bit-identical canonical set when unique, exact-valid always)."""
import ctypes

import numpy as np
import pytest

from gcdt_amd import DeviceBuffer, gen
from gcdt_amd.lib import check, lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VP = ctypes.c_void_p


def _ptr_release(p, nbytes):
    lib().gcdt_release(p, nbytes)


def run_reference_pipeline(g, cons, keep=None):
    """keep: a dict that receives host copies of the intermediate outputs."""
    L = lib()
    n = len(g)
    c = len(cons)
    pts = np.zeros((n + 4, 2), np.int32)
    pts[:n] = g
    d_pts = DeviceBuffer.from_array(pts)
    ms = ctypes.c_double(0)
    # getSamplePoints (rawMesh.h:6)
    sp, sid, ns = VP(), VP(), ctypes.c_int(0)
    check(L.gcdt_get_sample_points(None, 0, d_pts.ptr, n, ctypes.byref(sp), ctypes.byref(sid), ctypes.byref(ns),
                                   ctypes.byref(ms)), "getSamplePoints")
    S = ns.value
    if keep is not None:
        keep["sample_ids"] = DeviceBuffer.to_array(_View(sid.value), np.int32, (S,))
        keep["sample_pts"] = DeviceBuffer.to_array(_View(sp.value), np.int32, (S, 2))
    cap = max(4 * n + 4, 4 * S + 8)
    d_tris = DeviceBuffer(12 * cap)
    d_adj = DeviceBuffer(12 * cap)
    d_sons = DeviceBuffer(4 * cap)
    # rawMeshBefore (rawMesh.h:4) on the sample
    nt = ctypes.c_int(0)
    check(L.gcdt_raw_mesh_before(None, 0, sp.value, d_tris.ptr, d_sons.ptr, ctypes.byref(nt), S, ctypes.byref(ms)),
          "rawMeshBefore")
    if keep is not None:
        keep["before_pts"] = DeviceBuffer.to_array(_View(sp.value), np.int32, (S + 4, 2))
        keep["before_slots"] = nt.value
        keep["before_tris"] = d_tris.to_array(np.int32, (nt.value, 3))
        keep["before_sons"] = d_sons.to_array(np.int32, (nt.value,))
    # filterTriangles without adjacency
    d_ts = DeviceBuffer(12 * nt.value)
    check(L.gcdt_filter_triangles(None, 0, d_tris.ptr, d_ts.ptr, None, None, d_sons.ptr, ctypes.byref(nt),
                                  ctypes.byref(ms)), "filterTriangles")
    Ts = nt.value
    assert Ts == 2 * (S + 4) - 6
    # delaunaySample
    if keep is not None:
        keep["sample_leaves"] = d_ts.to_array(np.int32, (Ts, 3))
    check(L.gcdt_delaunay_sample(None, 0, sp.value, d_ts.ptr, d_adj.ptr, S + 4, Ts, ctypes.byref(ms)),
          "delaunaySample")
    if keep is not None:
        keep["sample_tris"] = d_ts.to_array(np.int32, (Ts, 3))
        keep["sample_adj"] = d_adj.to_array(np.int32, (Ts, 3))
    # getBitmap
    bo, bt = VP(), VP()
    check(L.gcdt_get_bitmap(None, 0, sp.value, d_ts.ptr, ctypes.byref(bo), ctypes.byref(bt), Ts, ctypes.byref(ms)),
          "getBitmap")
    nbt = np.zeros(1, np.int32)
    check(L.gcdt_memcpy_d2h(nbt.ctypes.data, bo.value + 4 * (1 << 20), 4), "d2h")
    if keep is not None:
        keep["bitmap_offsets"] = DeviceBuffer.to_array(_View(bo.value), np.int32, ((1 << 20) + 1,))
        keep["bitmap_tris"] = DeviceBuffer.to_array(_View(bt.value), np.int32, (int(nbt[0]),))
    # convertSample + copy into the main slots
    check(L.gcdt_convert_sample(d_ts.ptr, sid.value, S, n, Ts), "convertSample")
    check(L.gcdt_memcpy_d2d(d_tris.ptr, d_ts.ptr, 12 * Ts), "d2d")
    # rawMeshAfter (rawMesh.h:5)
    nt.value = Ts
    check(L.gcdt_raw_mesh_after(None, 0, d_pts.ptr, d_tris.ptr, d_adj.ptr, d_sons.ptr, bo.value, bt.value,
                                ctypes.byref(nt), n, ctypes.byref(ms)), "rawMeshAfter")
    slots = nt.value
    if keep is not None:
        keep["after_pts"] = d_pts.to_array(np.int32, (n + 4, 2))
        keep["after_slots"] = slots
        keep["after_tris"] = d_tris.to_array(np.int32, (slots, 3))
        keep["after_adj"] = d_adj.to_array(np.int32, (slots, 3))
        keep["after_sons"] = d_sons.to_array(np.int32, (slots,))
        # filterTriangles with adjacency on the unconstrained mesh
        f_t, f_a = DeviceBuffer(12 * slots), DeviceBuffer(12 * slots)
        nf = ctypes.c_int(slots)
        check(L.gcdt_filter_triangles(None, 0, d_tris.ptr, f_t.ptr, d_adj.ptr, f_a.ptr, d_sons.ptr,
                                      ctypes.byref(nf), None), "filterTriangles")
        keep["leaf_tris"] = f_t.to_array(np.int32, (nf.value, 3))
        keep["leaf_adj"] = f_a.to_array(np.int32, (nf.value, 3))
        f_t.free()
        f_a.free()
    # addConstraints (addConstraints.h:5)
    d_cons = DeviceBuffer.from_array(np.ascontiguousarray(cons, np.int32).reshape(-1, 2))
    ctris, nct = VP(), ctypes.c_int(0)
    check(L.gcdt_add_constraints(None, 0, d_pts.ptr, d_tris.ptr, d_adj.ptr, d_sons.ptr, bo.value, bt.value,
                                 d_cons.ptr if c else None, ctypes.byref(ctris), c, n, slots, ctypes.byref(nct),
                                 ctypes.byref(ms)), "addConstraints")
    # filterTriangles with adjacency, then append the constraint triangles
    d_new = DeviceBuffer(12 * slots)
    d_newadj = DeviceBuffer(12 * slots)
    nt.value = slots
    check(L.gcdt_filter_triangles(None, 0, d_tris.ptr, d_new.ptr, d_adj.ptr, d_newadj.ptr, d_sons.ptr,
                                  ctypes.byref(nt), ctypes.byref(ms)), "filterTriangles")
    T = nt.value
    if nct.value:
        check(L.gcdt_memcpy_d2d(d_new.ptr + 12 * T, ctris.value, 12 * nct.value), "d2d")
    T += nct.value
    assert T == 2 * n + 2
    # delaunayNew (kernel.cu:4697)
    pp, tp = VP(d_pts.ptr), VP(d_new.ptr)
    check(L.gcdt_delaunay_new(None, 0, ctypes.byref(pp), ctypes.byref(tp), d_cons.ptr if c else None, n + 4, T, c,
                              ctypes.byref(ms)), "delaunayNew")
    out = d_new.to_array(np.int32, (T, 3))
    # release library-owned buffers with their allocation sizes
    _ptr_release(sp.value, 8 * ((1 << 16) + 4))
    _ptr_release(sid.value, 4 * ((1 << 16) + 4))
    _ptr_release(bo.value, 4 * ((1 << 20) + 1))
    _ptr_release(bt.value, 4 * max(int(nbt[0]), 1))
    if nct.value:
        _ptr_release(ctris.value, 12 * nct.value)
    for b in (d_pts, d_tris, d_adj, d_sons, d_ts, d_cons, d_new, d_newadj):
        b.free()
    return out


class _View:
    """A library-owned device pointer read through DeviceBuffer.to_array."""

    def __init__(self, ptr):
        self.ptr = ptr


@pytest.mark.parametrize("case", ["uniform", "lattice"])
def test_reference_pipeline_intermediates(case):
    """The intermediate outputs of the drop-in pipeline against the oracle's
    restatements (oracle/dropin.py):
    * getSamplePoints: the smallest id of every occupied 256x256 cell, in
      cell order (rawMesh.cu:678-746), and their coordinates;
    * rawMeshBefore: the child layout of every split (rawMesh.cu:209-276);
      filterTriangles: the leaves in slot order;
    * delaunaySample: symmetric adjacency (testAdjTris, rawMesh.cu:931-992);
    * getBitmap: every cell's triangles equal the brute-force closed
      intersection (boostmap.cu:30-66) over each triangle's bounding cells;
    * rawMeshAfter: the child layout over all slots and symmetric adjacency
      among the leaves; filterTriangles with adjacency keeps it symmetric."""
    from oracle import dropin as D
    if case == "uniform":
        p, c = gen.uniform_cdt(n=30000, c=300)
    else:
        p, c = gen.lattice_cdt(k=96, c=80, max_len=40)
    g = gen.to_grid(p)
    k = {}
    tris = run_reference_pipeline(g, c, keep=k)
    assert O.validate(g, tris, c)["valid"]
    n = len(g)
    # sample
    want = D.sample_ids(g)
    assert np.array_equal(k["sample_ids"], want)
    assert np.array_equal(k["sample_pts"], g[want])
    S = len(want)
    # rawMeshBefore's insertion DAG and its leaves
    bp = k["before_pts"].astype(np.int64)
    assert np.array_equal(bp[:S], g[want]) and np.array_equal(bp[S:], O.with_corners(np.zeros((0, 2)))[:4])
    assert D.sons_violations(bp, k["before_tris"], k["before_sons"], k["before_slots"]) == []
    splits = int(np.sum(k["before_sons"] > 0) + np.sum(k["before_sons"] < -1) // 2)
    assert splits == S  # every sample point split exactly one triangle (or one edge)
    leaves = k["before_tris"][k["before_sons"] == -1]
    assert np.array_equal(k["sample_leaves"], leaves)
    # delaunaySample adjacency
    assert D.adj_violations(bp, k["sample_tris"], k["sample_adj"]) == []
    # bitmap CSR against brute force, cell by cell (order within a cell free)
    off, ids = D.bitmap_cells(bp, k["sample_tris"])
    go, gt = k["bitmap_offsets"].astype(np.int64), k["bitmap_tris"]
    assert np.array_equal(go, off)
    cell = np.repeat(np.arange(1 << 20), np.diff(go))
    got = gt[np.lexsort((gt, cell))]
    assert np.array_equal(got, ids)
    # rawMeshAfter's DAG and leaf adjacency
    ap = k["after_pts"].astype(np.int64)
    sons = k["after_sons"]
    assert D.sons_violations(ap, k["after_tris"], sons, k["after_slots"]) == []
    assert int(np.sum(sons > 0) + np.sum(sons < -1) // 2) == n - S
    assert D.adj_violations(ap, k["after_tris"], k["after_adj"], live=sons == -1) == []
    assert len(k["leaf_tris"]) == 2 * n + 2
    assert np.array_equal(k["leaf_tris"], k["after_tris"][sons == -1])
    assert D.adj_violations(ap, k["leaf_tris"], k["leaf_adj"]) == []


def test_solve_adjacency_output():
    """gcdt_solve's adjacency output is the rawMesh layout of the final mesh:
    symmetric, -1 only on the square's boundary (testAdjTris)."""
    from gcdt_amd import solve
    from oracle import dropin as D
    for p, c in (gen.uniform_cdt(n=20000, c=200), gen.lattice_cdt(k=64, c=40, max_len=30)):
        g = gen.to_grid(p)
        tris, st, adj = solve(g, c, with_adj=True)
        assert O.validate(g, tris, c)["valid"]
        assert D.adj_violations(O.with_corners(g), tris, adj) == []


@pytest.mark.parametrize("case", ["uniform", "lattice", "star"])
def test_reference_pipeline(case):
    if case == "uniform":
        p, c = gen.uniform_cdt(n=30000, c=300)
    elif case == "lattice":
        p, c = gen.lattice_cdt(k=96, c=80, max_len=40)
    else:
        p, c = gen.star_instance(gen.Stream(21), 20000)
    g = gen.to_grid(p)
    tris = run_reference_pipeline(g, c)
    v = O.validate(g, tris, c)
    assert v["valid"], v
    ref, _ = O.cdt(g, c)
    if O.validate(g, ref, c)["unique"]:
        assert np.array_equal(O.canonical(tris), O.canonical(ref))


def test_try_fast_delaunay_flips_any_triangulation():
    """fastDelaunay.h:5: any valid triangulation in, the Delaunay triangulation out."""
    L = lib()
    rng = np.random.default_rng(3)
    p, _ = gen.dedup(rng.integers(0, 1 << 20, size=(5000, 2)))
    g = gen.to_grid(p)
    n = len(g)
    pts = np.zeros((n + 4, 2), np.int32)
    pts[:n] = g
    d_pts = DeviceBuffer.from_array(pts)
    cap = 4 * n + 4
    d_tris, d_adj, d_sons = DeviceBuffer(12 * cap), DeviceBuffer(12 * cap), DeviceBuffer(4 * cap)
    nt = ctypes.c_int(0)
    ms = ctypes.c_double(0)
    # rawMesh (rawMesh.h:7): insertion only, no flips -> a non-Delaunay triangulation
    check(L.gcdt_raw_mesh(None, 0, d_pts.ptr, d_tris.ptr, d_adj.ptr, d_sons.ptr, ctypes.byref(nt), n, None), "rawMesh")
    d_leaf, d_leafadj = DeviceBuffer(12 * nt.value), DeviceBuffer(12 * nt.value)
    check(L.gcdt_filter_triangles(None, 0, d_tris.ptr, d_leaf.ptr, d_adj.ptr, d_leafadj.ptr, d_sons.ptr,
                                  ctypes.byref(nt), None), "filter")
    T = nt.value
    before = d_leaf.to_array(np.int32, (T, 3))
    vb = O.validate(g, before, None)
    assert vb["not_delaunay"] > 0 and vb["not_ccw"] == 0 and vb["edge_mismatch"] == 0
    check(L.gcdt_try_fast_delaunay(None, 0, d_pts.ptr, d_leaf.ptr, d_leafadj.ptr, n + 4, T, ctypes.byref(ms)),
          "tryFastDelaunay")
    after = d_leaf.to_array(np.int32, (T, 3))
    ref, _ = O.cdt(g, None)
    assert np.array_equal(O.canonical(after), O.canonical(ref))
    for b in (d_pts, d_tris, d_adj, d_sons, d_leaf, d_leafadj):
        b.free()


def test_scratch_too_small_is_reported():
    L = lib()
    g = gen.to_grid(gen.uniform_cdt(n=3000, c=0)[0])
    d = DeviceBuffer.from_array(g)
    tmp = DeviceBuffer(8)
    sp, sid, ns = VP(), VP(), ctypes.c_int(0)
    rc = L.gcdt_get_sample_points(tmp.ptr, 8, d.ptr, len(g), ctypes.byref(sp), ctypes.byref(sid), ctypes.byref(ns),
                                  None)
    assert rc == 6  # GCDT_ERR_SCRATCH
    d.free()
    tmp.free()
