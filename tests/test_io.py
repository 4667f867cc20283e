"""Input/output formats: gcdt_read_file (the reference's readFile,
kernel.cu:5159-5291), gcdt_canonicalize_host and gcdt_write_mesh run on the
host, so these tests need no GPU; gcdt_trans_points is in test_gpu_io.py."""
import numpy as np
import pytest

from gcdt_amd import GcdtError
from gcdt_amd import io as gio
from oracle import io_oracle as IO
from oracle import oracle as O

KAT = """5
0 0 0
1 10 0
2 10 5
3 0 0
4 5 10
5
0 1 2
1 2 1
2 2 2
3 4 3
4 0 4
"""


def _write(tmp_path, text, name="in.txt"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_known_answer_file(tmp_path):
    """Hand-worked: file idx 3 repeats idx 0, so 4 points remain; the box is
    [0,10]^2 (centre 5, side 10), so 0 -> 0.005, 5 -> 0.5, 10 -> 0.995.
    Constraints ("k a b"): (1,2) kept; (2,1) repeats it reversed -> dropped;
    (2,2) is a self-loop -> dropped; file ids (4,3) -> slots (3,0) kept;
    (0,4) -> (0,3) repeats that reversed -> dropped."""
    pts, cons = gio.read_file(_write(tmp_path, KAT))
    want = np.array([[0.005, 0.005], [0.995, 0.005], [0.995, 0.5], [0.5, 0.995]])
    assert pts.shape == (4, 2)
    assert np.allclose(pts, want, rtol=0, atol=1e-15)
    assert pts[2, 1] == 0.5 and pts[3, 0] == 0.5   # the centre maps exactly
    assert cons.tolist() == [[1, 2], [3, 0]]
    opts, ocons = IO.read_file(KAT)
    assert np.array_equal(opts, pts) and np.array_equal(ocons, cons)


def _random_text(rng, n, c, dup_frac=0.2, scale=1000.0, offset=0.0):
    base = rng.uniform(-scale, scale, size=(n, 2)) + offset
    k = int(n * dup_frac)
    base[rng.integers(0, n, k)] = base[rng.integers(0, n, k)]      # exact duplicates
    ids = rng.permutation(n)                                          # ids listed out of order
    lines = [str(n)] + [f"{i} {float(base[i, 0])!r} {float(base[i, 1])!r}" for i in ids]
    a = rng.integers(0, n, c)
    b = rng.integers(0, n, c)
    b[: c // 10] = a[: c // 10]                                       # self-loops
    lines.append(str(c))
    lines += [f"{j} {int(a[j])} {int(b[j])}" for j in range(c)]
    lines += [f"{c + j} {int(b[j])} {int(a[j])}" for j in range(0)]
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("seed,n,c", [(1, 50, 20), (2, 3000, 800), (3, 1, 0), (4, 2, 3)])
def test_read_file_matches_oracle(tmp_path, seed, n, c):
    rng = np.random.default_rng(seed)
    text = _random_text(rng, n, c)
    pts, cons = gio.read_file(_write(tmp_path, text))
    opts, ocons = IO.read_file(text)
    assert np.array_equal(pts, opts)
    assert np.array_equal(cons, ocons)
    if len(pts):
        assert pts.min() >= 0.005 - 1e-12 and pts.max() <= 0.995 + 1e-12


def test_far_coordinates_use_the_true_bounding_box(tmp_path):
    """Every coordinate above 1e6 (where the reference's +-1e6 box start is
    wrong): the longer side still spans exactly [0.005, 0.995]."""
    text = "3\n0 2000000 3000000\n1 2000010 3000000\n2 2000000 3000004\n0\n"
    pts, _ = gio.read_file(_write(tmp_path, text))
    assert np.isclose(pts[:, 0].min(), 0.005) and np.isclose(pts[:, 0].max(), 0.995)


def test_shuffle_is_a_permutation_with_constraints_following(tmp_path):
    rng = np.random.default_rng(9)
    text = _random_text(rng, 2000, 500)
    path = _write(tmp_path, text)
    plain, pc = gio.read_file(path)
    shuf, sc = gio.read_file(path, shuffle=True)
    assert shuf.shape == plain.shape
    # recover the permutation from the coordinates (unique after the merge)
    where = {tuple(p): i for i, p in enumerate(shuf)}
    perm = np.array([where[tuple(p)] for p in plain])
    assert sorted(perm.tolist()) == list(range(len(plain)))
    assert not np.array_equal(perm, np.arange(len(plain)))
    assert np.array_equal(sc, perm[pc])
    opts, ocons = IO.read_file(text, perm=perm)
    assert np.array_equal(shuf, opts) and np.array_equal(sc, ocons)
    # the same file shuffles the same way every time
    again, _ = gio.read_file(path, shuffle=True)
    assert np.array_equal(again, shuf)


def test_oracle_planted_mistake_is_caught(tmp_path):
    """A restatement that kept the LAST duplicate instead of the first would
    move a point; the known-answer file distinguishes them."""
    text = "3\n0 0 0\n1 10 10\n2 0 0\n1\n0 2 1\n"
    pts, cons = gio.read_file(_write(tmp_path, text))
    assert cons.tolist() == [[0, 1]]      # file id 2 merges into slot 0, not a new slot 2
    assert len(pts) == 2


@pytest.mark.parametrize("text,msg", [
    ("3\n0 0 0\n5 1 1\n2 2 2\n", "out of range"),
    ("2\n0 0 0\n", "truncated"),
    ("2\n0 0 0\n1 1 1\n1\n0 0 7\n", "not a listed point"),
    ("2\n0 0 0\n1 nan 1\n", "non-finite"),
])
def test_malformed_files_fail(tmp_path, text, msg):
    with pytest.raises(GcdtError, match=msg):
        gio.read_file(_write(tmp_path, text))


def test_missing_file_fails(tmp_path):
    with pytest.raises(GcdtError, match="cannot open"):
        gio.read_file(str(tmp_path / "nope.txt"))


def test_canonical_matches_oracle():
    rng = np.random.default_rng(5)
    t = rng.integers(0, 1000, size=(5000, 3)).astype(np.int32)
    assert np.array_equal(gio.canonical(t), O.canonical(t))
    # rotation keeps orientation: (2,0,1) -> (0,1,2), never (0,2,1)
    assert gio.canonical(np.array([[2, 0, 1], [5, 9, 3]], np.int32)).tolist() == [[0, 1, 2], [3, 5, 9]]


def test_write_mesh_round_trip(tmp_path):
    pts = np.array([[1, 2], [30, 40], [5, 600]], np.int32)
    tris = np.array([[0, 1, 2], [2, 1, 0]], np.int32)
    path = str(tmp_path / "out.txt")
    gio.write_mesh(path, pts, tris)
    tok = open(path).read().split()
    assert int(tok[0]) == 3
    body = np.array(tok[1:10], np.int64).reshape(3, 3)
    assert np.array_equal(body[:, 1:], pts) and body[:, 0].tolist() == [0, 1, 2]
    assert int(tok[10]) == 2
    assert np.array_equal(np.array(tok[11:], np.int64).reshape(2, 4)[:, 1:], tris)


def test_write_input_round_trips_through_read_file(tmp_path):
    rng = np.random.default_rng(12)
    p = rng.uniform(0, 1, size=(100, 2))
    path = str(tmp_path / "f.txt")
    gio.write_input(path, p, np.array([[0, 1], [1, 2]]))
    pts, cons = gio.read_file(path)
    ref, rc = IO.read_file(open(path).read())
    assert np.array_equal(pts, ref) and cons.tolist() == [[0, 1], [1, 2]]
