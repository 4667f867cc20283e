"""GPU side of the input path: gcdt_trans_points against the oracle
(bit-exact), and a file -> grid -> CDT -> canonical list run that matches the
oracle's pipeline on the same file."""
import numpy as np
import pytest

from gcdt_amd import gen, solve
from gcdt_amd import io as gio
from oracle import io_oracle as IO
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_trans_points_bit_exact():
    rng = np.random.default_rng(4)
    f = rng.uniform(0, 1, size=(200_000, 2))
    edge = np.array([[0.0, 0.0], [1.0, 1.0], [-0.25, 1.5], [0.999999999999, 1e-300],
                     [np.nextafter(1.0, 0.0), 0.5], [0.005, 0.995]])
    f = np.concatenate([edge, f])
    g = gio.trans_points(f)
    assert np.array_equal(g, IO.trans_points(f))
    assert g[1].tolist() == [(1 << 30) - 1, (1 << 30) - 1] and g[2].tolist() == [0, (1 << 30) - 1]


def test_file_to_canonical_mesh(tmp_path):
    s = gen.Stream(21)
    p, c = gen.polygon_instance(s, 20_000)
    pf = (p.astype(np.float64) + 0.25) / float(1 << 20)
    path = str(tmp_path / "inst.txt")
    gio.write_input(path, pf, c)
    fpts, cons = gio.read_file(path)
    g = gio.trans_points(fpts)
    tris, _ = solve(g, cons)
    out = gio.canonical(tris)
    # the oracle's pipeline on the same file
    ofp, ocons = IO.read_file(open(path).read())
    og = IO.trans_points(ofp)
    assert np.array_equal(g, og) and np.array_equal(cons, ocons)
    ref, _ = O.cdt(og, ocons)
    v = O.validate(og, tris, ocons)
    assert v["valid"], v
    if v["unique"]:
        assert np.array_equal(out, O.canonical(ref))
    gio.write_mesh(str(tmp_path / "mesh.txt"), g, out)


def test_reference_four_point_case_through_the_input_path():
    """kernel.cu:49-52 points and the chain of kernel.cu:5348-5351, mapped on
    the GPU by gcdt_trans_points (kernel.cu:171-188) and solved: the result is
    the brute-force CDT."""
    f = np.array([[0.75, 0.5], [0.75, 0.2], [0.25, 0.2], [0.5, 0.75]])
    g = gio.trans_points(f)
    assert np.array_equal(g, IO.trans_points(f))
    cons = np.array([[0, 1], [1, 2], [2, 3]], np.int32)
    tris, _ = solve(g, cons)
    assert np.array_equal(gio.canonical(tris), O.brute_force_cdt(g, cons))
