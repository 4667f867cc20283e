"""The library's kernels and round drivers executed on the CPU by the SIMT
emulator (tests/emu): every device allocation sits between guard pages, new
allocations hold random bytes, workgroups run interleaved and work-items
switch at every atomic, so an out-of-bounds list write, a read of an unwritten
entry used as an index, or a pool release with the wrong size ends the child
process with the kernel and work-item named.

Covered: every value of every round tunable that the autotuner sweeps (and
grid caps far below the GPU's, so grid-stride loops run many iterations),
on mixed-size batches, single solves with constraints, and the degenerate
lattice.  Each result must also pass the oracle's exact validator."""
import json
import os
import subprocess
import sys
import zlib

import pytest

from tests.emu import emu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SWEEP = {
    "tail_threshold": [0, 16, 512, 1024, 2048, 8192, 16384, 1 << 20],
    "grid_cap": [1, 3, 2048, 16384],
    "flip_items": [2, 4],
    "ins_items": [2, 4],
    "pivot": [0],
    "single_tester": [0],
}
BATCH = [64, 300, 2000, 900, 150, 5000, 70, 12000]


@pytest.fixture(scope="module", autouse=True)
def _built():
    emu.build()


def _run(case, **kw):
    """yield_mem=2 also switches work-items between the 16-byte parts of
    record loads and stores, stalling some for many passes, so records are
    read while flips rewrite them."""
    rc, out, err = emu.run_in_child(case, **kw)
    assert rc == 0, f"emulated run failed with status {rc}:\n{err}"
    assert out is not None and out["valid"], out
    assert out["unique_mismatch"] == 0, out
    return out


@pytest.mark.parametrize("knob,value", [(k, v) for k, vs in SWEEP.items() for v in vs])
def test_batch_at_every_swept_value(knob, value):
    _run({"kind": "batch", "sizes": BATCH, "seed": 7, "knobs": {knob: value}},
         seed=zlib.crc32(f"{knob}={value}".encode()) % 1000 + 1, align_low=value & 1, conc=8, yield_atomics=1, yield_mem=2)


def test_batch_all_knobs_off_default():
    knobs = {"tail_threshold": 64, "grid_cap": 3, "flip_items": 2, "ins_items": 4, "pivot": 0, "single_tester": 0}
    for seed in (11, 12):
        _run({"kind": "batch", "sizes": BATCH, "seed": seed, "knobs": knobs}, seed=seed, align_low=seed & 1, conc=16,
             yield_atomics=1)


def test_star_instances():
    _run({"kind": "batch", "count": 6, "nmax": 20000, "seed": 6}, seed=5, conc=8, yield_atomics=1)


@pytest.mark.parametrize("tail", [0, 4096, 1 << 20])
def test_single_solve_with_constraints(tail):
    _run({"kind": "uniform", "n": 20000, "c": 200, "knobs": {"tail_threshold": tail}}, seed=tail % 97 + 1,
         align_low=1, conc=8, yield_atomics=1)


def test_lattice_degenerate():
    _run({"kind": "lattice", "k": 48, "c": 40, "max_len": 20, "knobs": {"grid_cap": 2}}, seed=3, conc=4,
         yield_atomics=1)


def test_clustered_fans():
    _run({"kind": "clustered", "n": 30000, "circles": 10}, seed=9, conc=8, yield_atomics=1)


@pytest.mark.parametrize("where", ["past_end", "before_start"])
def test_detector_catches_an_out_of_bounds_write(where):
    """The guard pages do fault: a write one element past the end (or before
    the start) of an allocation ends the child with status 86 and names the
    allocation and the offset."""
    code = ("import ctypes; from tests.emu import emu; L = emu.load(); "
            f"L.emu_selftest_oob.argtypes = [ctypes.c_int]; L.emu_selftest_oob({1 if where == 'past_end' else -1})")
    env = dict(os.environ, EMU_ALIGN_LOW="0" if where == "past_end" else "1")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 86, (r.returncode, r.stderr)
    assert "emu_selftest" in r.stderr and "offset" in r.stderr, r.stderr


def test_detector_catches_a_wrong_size_release():
    code = "from tests.emu import emu; L = emu.load(); L.emu_selftest_release()"
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 91, (r.returncode, r.stderr)


def test_noalloc_solve_equals_solve_on_leftovers():
    """gcdt_solve_noalloc (every buffer carved from the static scratch array)
    against gcdt_solve on three inputs in a row, the later ones smaller, so
    they run over the earlier calls' contents of the array."""
    out = _run({"kind": "noalloc", "inputs": [["uniform", 12000, 120, 51], ["star", 5000, 52],
                                              ["lattice", 40, 20, 12]]},
               seed=4, conc=8, yield_atomics=1)
    assert out["status"] == [0, 0, 0], out


def test_noalloc_solve_refuses_without_setup_and_above_its_maximum():
    """Without gcdt_pool_init on the device there is no scratch address: the
    call fails with GCDT_ERR_INVALID instead of touching the array; and an
    input above GCDT_NOALLOC_MAX_POINTS fails with GCDT_ERR_SCRATCH."""
    rc, out, err = emu.run_in_child({"kind": "noalloc", "init": False, "inputs": [["uniform", 1000, 10, 53]]})
    assert rc == 0 and out["status"] == [2], (rc, out, err)
    rc, out, err = emu.run_in_child({"kind": "noalloc", "inputs": [["uniform", 800000, 0, 54]]})
    assert rc == 0 and out["status"] == [6], (rc, out, err)


# ---------------------------------------------------------------- crossing constraints

@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_steiner_tiny_inputs_equal_the_brute_force_cdt(seed):
    """Crossing pieces deferred by the cavity rounds (workgroups interleaved
    by the emulator seed) and inserted by the serial pass: the exact
    validator, and the brute-force CDT with the crossing vertices."""
    out = _run({"kind": "steiner", "gen": ["random", seed, 12, 6], "brute": True}, seed=seed, conc=8,
               yield_atomics=1)
    assert out.get("status") is None


@pytest.mark.parametrize("gen", [["star", 6, 4, 5], ["hd", 8, 5, 6, 40], ["random", 9, 400, 40]])
def test_steiner_star_grid_crossings_and_many_crossings(gen):
    """Four segments through one point; 30 grid crossings equal to the
    oracle's CDT of the points plus the crossings; hundreds of crossings."""
    out = _run({"kind": "steiner", "gen": gen, "augmented": gen[0] == "hd", "max_steiner": 2000}, seed=3, conc=8,
               yield_atomics=1)
    assert out["steiner"] > 0


def test_steiner_capacity_is_reported():
    rc, out, err = emu.run_in_child({"kind": "steiner", "gen": ["random", 9, 400, 40], "max_steiner": 3})
    assert rc == 0 and out["status"] == 8, (out, err)


# ---------------------------------------------------------------- determinism

@pytest.mark.parametrize("case", [{"kind": "lattice", "k": 80, "c": 12, "max_len": 20},
                                  {"kind": "clustered", "n": 12000, "circles": 6},
                                  {"kind": "batch", "sizes": [64, 300, 2000, 900, 150, 20000], "seed": 7},
                                  {"kind": "steiner", "gen": ["random", 9, 400, 40], "max_steiner": 2000}],
                         ids=["lattice", "clustered", "batch", "crossings"])
def test_same_output_under_every_interleaving(case):
    """Triangle and adjacency arrays are bit-identical whatever order the
    emulated workgroups and work-items run in (atomics arrive in a different
    order each time), degenerate lattice and batches included."""
    digests = set()
    for seed, conc in ((1, 1), (2, 8), (3, 16)):
        out = _run(dict(case, digest=1), seed=seed, conc=conc, yield_atomics=1)
        digests.add(out["digest"])
    assert len(digests) == 1


def test_per_cell_reorder_paths_keep_valid_meshes():
    """With the per-cell reorder's caps at 3 and 8 elements, the lane-rank,
    the workgroup and the large-cell merge paths all run on small meshes, and
    with the wave cavity cap at 4 most cavities are retriangulated by a wave:
    every result must still pass the validator (a reorder that lost or
    doubled an element would corrupt the point or leaf order; a wrong
    Delaunay vertex would leave a non-Delaunay edge)."""
    lib = emu.build(emu.CELLCAP)
    for case in ({"kind": "lattice", "k": 64, "c": 8, "max_len": 12}, {"kind": "uniform", "n": 6000, "c": 40},
                 {"kind": "batch", "sizes": [64, 700, 300], "seed": 9}):
        rc, out, err = emu.run_in_child(case, seed=5, conc=8, yield_atomics=1, lib=lib)
        assert rc == 0 and out is not None and out["valid"], (case, out, err)


@pytest.mark.parametrize("case", [{"kind": "lattice", "k": 64, "c": 8, "max_len": 12},
                                  {"kind": "clustered", "n": 4000, "circles": 3},
                                  {"kind": "batch", "sizes": [64, 700, 300], "seed": 9}],
                         ids=["lattice", "clustered", "batch"])
def test_large_cells_are_deterministic_too(case):
    """With the workgroup cap at 8 elements most cells take the large-cell
    merge path; the arrays must still be the same under every interleaving
    (cells above the cap once kept their arrival order)."""
    lib = emu.build(emu.CELLCAP)
    digests = set()
    for seed, conc in ((1, 1), (2, 8), (3, 16)):
        rc, out, err = emu.run_in_child(dict(case, digest=1), seed=seed, conc=conc, yield_atomics=1, lib=lib)
        assert rc == 0 and out is not None and out["valid"], (case, out, err)
        digests.add(out["digest"])
    assert len(digests) == 1


@pytest.mark.parametrize("case", [{"kind": "lattice", "k": 64, "c": 8, "max_len": 12},
                                  {"kind": "uniform", "n": 6000, "c": 40},
                                  {"kind": "batch", "sizes": [64, 700, 300], "seed": 9}],
                         ids=["lattice", "uniform", "batch"])
def test_small_cap_paths_give_the_default_arrays(case):
    """Every path of the per-cell reorder places the same order, and the wave
    retriangulation picks the same Delaunay vertices as one thread's scan
    (cocircular lattice included): the small-cap build's arrays equal the
    default build's, bit for bit."""
    small = emu.build(emu.CELLCAP)
    digests = []
    for lib in (None, small):
        rc, out, err = emu.run_in_child(dict(case, digest=1), seed=2, conc=8, yield_atomics=1, lib=lib)
        assert rc == 0 and out is not None and out["valid"], (case, out, err)
        digests.append(out["digest"])
    assert digests[0] == digests[1]


# ---------------------------------------------------------------- phase checks

def test_mesh_check_catches_planted_mistakes():
    """gcdt_check_mesh: a solved Delaunay mesh passes; one triangle turned
    clockwise, one wrong or out-of-range neighbour, and one flipped convex
    edge (valid, counter-clockwise, not locally Delaunay) are each caught with
    their check bits and the lowest failing triangle."""
    rc, out, err = emu.run_in_child({"kind": "check", "n": 300}, seed=1, conc=4, yield_atomics=1)
    assert rc == 0 and out is not None, err
    assert out["clean"] == [0, -1]
    assert out["swapped"][0] & 2 and out["swapped"][0] & 4 and out["swapped"][1] <= 5
    assert out["adjacency"][0] == 4 and out["adjacency"][1] <= 7
    assert out["range"] == [4, 9]
    assert out["nondelaunay"][0] == 16 and out["nondelaunay_without_check"] == [0, -1]


@pytest.mark.parametrize("case", [{"kind": "uniform", "n": 6000, "c": 40},
                                  {"kind": "lattice", "k": 48, "c": 20, "max_len": 12},
                                  {"kind": "clustered", "n": 8000, "circles": 4},
                                  {"kind": "batch", "sizes": [64, 700, 300, 2000], "seed": 9}],
                         ids=["uniform", "lattice", "clustered", "batch"])
def test_solves_pass_the_phase_checks(case):
    """With the validate tunable on, the mesh is checked on the device after
    every phase (sample mesh, its flips, the flips, the constraints); a
    failed check would fail the solve."""
    _run(dict(case, knobs={"validate": 1}), seed=6, conc=8, yield_atomics=1)
