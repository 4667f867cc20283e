"""Seeded synthetic CDT workloads (points + constraint segments).

All randomness comes from SplitMix64 streams; no STL/numpy distribution is
used (the reference's generators, kernel.cu:43-170, rely on
implementation-defined std::default_random_engine streams and cannot be
reproduced).  The point sets of the uniform, clustered and lattice workloads
and every polygon built from circles are computed in integer arithmetic only
(Irwin-Hall Gaussian offsets, fixed-point cos/sin), so they are the same on
every machine.  The star instances' vertex angles and the log-uniform sizes and
Poisson arrival times of the batched and serving workloads use double
arithmetic and the C library's cos/sin/log/exp, rounded to integers; a last-bit
difference between C libraries could move a value that lies within one ulp of
a rounding boundary.  Coordinates are integers in [0, 2^20); :func:`to_grid` maps them
exactly into the reference's 2^30 point grid (kernel.cu:171-188 maps doubles
into the same grid) by x' = (x << 10) + 512, a positive scale plus a
translation, so every orientation / in-circle sign is preserved and every
input lies strictly inside the bounding square (0, 2^30-1)^2.

Workloads:
  uniform_cdt(n, c)       points uniform + c horizontal segments
  clustered_cdt(n, c)     80% uniform / 20% Gaussian clusters + 100-gon circles
  lattice_cdt(k, c)       k x k lattice + +-45 degree diagonal segments
  star_instance(n, ...)   star-shaped outer polygon with star-shaped holes
  batched_instances(...)  many star instances, log-uniform sizes
  serving_requests(...)   Poisson arrivals of uniform+polygon instances
"""
from __future__ import annotations

import math

import numpy as np

COORD_BITS = 20
COORD_MAX = 1 << COORD_BITS
GRID_BITS = 30
GRID_R = (1 << GRID_BITS) - 1

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class Stream:
    """SplitMix64 counter-mode stream (Steele et al. 2014): value i is
    mix(seed + (i+1) * golden)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.pos = 0

    def u64(self, n: int) -> np.ndarray:
        i = np.arange(self.pos + 1, self.pos + n + 1, dtype=np.uint64)
        self.pos += n
        with np.errstate(over="ignore"):
            z = self.seed + i * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))

    def bits(self, n: int, b: int) -> np.ndarray:
        """n values uniform in [0, 2^b)."""
        return (self.u64(n) >> np.uint64(64 - b)).astype(np.int64)

    def below(self, n: int, m: int) -> np.ndarray:
        """n values in [0, m) (multiply-shift, bias < 2^-32)."""
        hi = self.u64(n) >> np.uint64(32)
        return ((hi.astype(np.uint64) * np.uint64(m)) >> np.uint64(32)).astype(np.int64)

    def unit(self, n: int) -> np.ndarray:
        """n doubles in [0, 1) with 53 random bits."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def to_grid(xy: np.ndarray) -> np.ndarray:
    """[0,2^20)^2 integer coordinates -> the 2^30 grid, exactly."""
    xy = np.asarray(xy, dtype=np.int64)
    return ((xy << (GRID_BITS - COORD_BITS)) + (1 << (GRID_BITS - COORD_BITS - 1))).astype(np.int32)


def dedup(xy: np.ndarray):
    """Drop repeated points, keeping first occurrences in order.
    Returns (unique points, map old index -> new index)."""
    xy = np.asarray(xy, dtype=np.int64).reshape(-1, 2)
    key = (xy[:, 0] << COORD_BITS) | xy[:, 1]
    _, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    newid = np.empty(len(first), dtype=np.int64)
    newid[order] = np.arange(len(first))
    return xy[first[order]].astype(np.int32), newid[inv.reshape(-1)]


def _finish(pts_list, seg_pts):
    """Concatenate point blocks and segment endpoints, dedup, and return
    (points int32 (n,2), constraints int32 (c,2) as vertex ids)."""
    blocks = [np.asarray(p, dtype=np.int64).reshape(-1, 2) for p in pts_list]
    npts = sum(len(b) for b in blocks)
    segs = np.asarray(seg_pts, dtype=np.int64).reshape(-1, 2, 2)
    allp = np.concatenate(blocks + [segs.reshape(-1, 2)]) if len(segs) else np.concatenate(blocks)
    uniq, remap = dedup(allp)
    if len(segs):
        ids = remap[npts:].reshape(-1, 2)
        keep = ids[:, 0] != ids[:, 1]
        cons = ids[keep].astype(np.int32)
    else:
        cons = np.zeros((0, 2), dtype=np.int32)
    return uniq, cons


def uniform_cdt(n: int = 100_000, c: int = 1_000, seed_pts: int = 1, seed_cons: int = 2):
    """Plain CDT, small: n uniform points, c horizontal segments at c distinct
    y values with random x endpoints (non-crossing by construction)."""
    s = Stream(seed_pts)
    pts = np.stack([s.bits(n, COORD_BITS), s.bits(n, COORD_BITS)], axis=1)
    s2 = Stream(seed_cons)
    ys = []
    seen = set()
    while len(ys) < c:
        for y in s2.bits(2 * c, COORD_BITS).tolist():
            if y not in seen:
                seen.add(y)
                ys.append(y)
                if len(ys) == c:
                    break
    segs = []
    for y in ys:
        while True:
            x1, x2 = s2.bits(2, COORD_BITS).tolist()
            if x1 != x2:
                break
        if x1 > x2:
            x1, x2 = x2, x1
        segs.append(((x1, y), (x2, y)))
    return _finish([pts], segs)


# Fixed-point trigonometry in Python integers (exact and the same on every
# machine, unlike a C library's cos/sin): values scaled by 2^TRIG_BITS.
TRIG_BITS = 40
_WORK = 96  # working precision of the series


def _atan_inv(x: int) -> int:
    """atan(1/x) * 2^_WORK by its alternating series."""
    one = 1 << _WORK
    term = one // x
    total, k, x2 = 0, 0, x * x
    while term:
        total += term // (2 * k + 1) if k % 2 == 0 else -(term // (2 * k + 1))
        term //= x2
        k += 1
    return total


_PI = 16 * _atan_inv(5) - 4 * _atan_inv(239)  # Machin's formula, pi * 2^_WORK


def cos_sin_fixed(k: int, m: int):
    """(cos, sin) of 2*pi*k/m, each round(v * 2^TRIG_BITS), from Taylor series
    in integer arithmetic."""
    one = 1 << _WORK
    k %= m
    if 2 * k > m:  # angle in (-pi, pi]
        k -= m
    a = (2 * _PI * k) // m
    c, s_, term, n = 0, 0, one, 0
    while term:
        if n % 2 == 0:
            c += term if n % 4 == 0 else -term
        else:
            s_ += term if n % 4 == 1 else -term
        n += 1
        q = abs(term * a) // (n << _WORK)  # truncated towards zero, so the series ends
        term = q if term * a >= 0 else -q
    sh = _WORK - TRIG_BITS
    half = 1 << (sh - 1)
    return (c + half) >> sh, (s_ + half) >> sh


def _circle_segments(s: Stream, count: int, k: int, rmin: int, rmax: int, taken=None):
    """count disjoint k-gons with pairwise disjoint bounding boxes, all in
    integer arithmetic: integer centre and radius, vertex j at
    centre + round(r * (cos, sin)(2*pi*j/k)) from cos_sin_fixed.  Returns a
    list of ((x1,y1),(x2,y2)) segments."""
    boxes = [] if taken is None else taken
    segs = []
    tries = 0
    cs = np.array([cos_sin_fixed(j, k) for j in range(k)], dtype=np.int64)
    half = 1 << (TRIG_BITS - 1)
    while len(boxes) < count and tries < 200 * count + 1000:
        tries += 1
        r = int(rmin + s.below(1, rmax - rmin + 1)[0])
        cx, cy = (s.below(2, COORD_MAX - 2 * r - 4) + r + 2).tolist()
        box = (cx - r - 2, cy - r - 2, cx + r + 2, cy + r + 2)
        if any(not (box[2] < b[0] or b[2] < box[0] or box[3] < b[1] or b[3] < box[1]) for b in boxes):
            continue
        boxes.append(box)
        vx = cx + ((r * cs[:, 0] + half) >> TRIG_BITS)
        vy = cy + ((r * cs[:, 1] + half) >> TRIG_BITS)
        for i in range(k):
            j = (i + 1) % k
            segs.append(((int(vx[i]), int(vy[i])), (int(vx[j]), int(vy[j]))))
    return segs


def gaussian_offsets(s: Stream, n: int, log2_sigma: int = 12) -> np.ndarray:
    """n integer offsets, approximately normal with standard deviation
    2^log2_sigma, in integer arithmetic only: the sum of 12 uniform 16-bit
    integers (variance 12 * (2^32 - 1) / 12, about 2^32) centred and scaled
    by 2^(log2_sigma - 16) with rounding (the Irwin-Hall construction; its
    tails end at 6 sigma)."""
    acc = np.zeros(n, dtype=np.int64)
    for _ in range(12):
        acc += s.bits(n, 16)
    g = acc - 12 * ((1 << 16) - 1) // 2
    sh = 16 - log2_sigma
    return (g + (1 << (sh - 1))) >> sh


def clustered_cdt(n: int = 10_000_000, circles: int = 1_000, k: int = 100, seed_pts: int = 3,
                  seed_cons: int = 4, log2_sigma: int = 12, centres: int = 64):
    """Plain CDT, large: 80% uniform / 20% Gaussian-cluster points (sigma
    2^12 around 64 uniform centres, clamped to the coordinate range) plus
    `circles` disjoint k-gons (radius >= 2^9) as constraints.  Integer
    arithmetic throughout."""
    s = Stream(seed_pts)
    cen = s.bits(2 * centres, COORD_BITS).reshape(-1, 2)
    clustered = s.below(n, 5) == 4
    ux = s.bits(n, COORD_BITS)
    uy = s.bits(n, COORD_BITS)
    cid = s.below(n, centres)
    gx = np.clip(cen[cid, 0] + gaussian_offsets(s, n, log2_sigma), 0, COORD_MAX - 1)
    gy = np.clip(cen[cid, 1] + gaussian_offsets(s, n, log2_sigma), 0, COORD_MAX - 1)
    x = np.where(clustered, gx, ux)
    y = np.where(clustered, gy, uy)
    s2 = Stream(seed_cons)
    rmax = min(4096, COORD_MAX // (4 * math.isqrt(max(circles, 1))))
    segs = _circle_segments(s2, circles, k, 512, max(rmax, 600))
    return _finish([np.stack([x, y], axis=1)], segs)


def lattice_cdt(k: int = 4096, c: int = 2_000, step: int = 256, seed_cons: int = 5, max_len: int = 512):
    """Degenerate stress: k x k lattice (spacing `step`) plus c diagonal
    segments between lattice points, lengths in [1, max_len] lattice steps,
    no two sharing a lattice cell (so none overlap or cross)."""
    i, j = np.meshgrid(np.arange(k, dtype=np.int64), np.arange(k, dtype=np.int64), indexing="ij")
    pts = np.stack([i.reshape(-1) * step, j.reshape(-1) * step], axis=1)
    s = Stream(seed_cons)
    used = set()
    cons = []
    tries = 0
    max_len = max(1, min(max_len, k - 1))
    while len(cons) < c and tries < 100 * c + 100:
        tries += 1
        L = 1 + int(s.below(1, max_len)[0])
        up = int(s.below(1, 2)[0])
        if k - L <= 0:
            continue
        x0 = int(s.below(1, k - L)[0])
        y0 = int(s.below(1, k - L)[0]) + (L if not up else 0)
        dy = 1 if up else -1
        cells = [(x0 + t, min(y0 + t * dy, y0 + (t + 1) * dy)) for t in range(L)]
        if any(cc in used for cc in cells):
            continue
        used.update(cells)
        a = x0 * k + y0
        b = (x0 + L) * k + (y0 + L * dy)
        cons.append((a, b))
    return pts.astype(np.int32), np.asarray(cons, dtype=np.int32).reshape(-1, 2)


def _star_polygon(s: Stream, cx, cy, r, m):
    ang = (np.arange(m) + 0.5 * s.unit(m)) * (2.0 * math.pi / m)
    rad = r * (0.7 + 0.3 * s.unit(m))
    vx = np.rint(cx + rad * np.cos(ang)).astype(np.int64)
    vy = np.rint(cy + rad * np.sin(ang)).astype(np.int64)
    return [((int(vx[i]), int(vy[i])), (int(vx[(i + 1) % m]), int(vy[(i + 1) % m]))) for i in range(m)]


def star_instance(s: Stream, n: int):
    """One instance: star-shaped outer polygon (16-512 vertices), 0-4
    star-shaped holes, uniform interior points; n points in total."""
    cx = cy = COORD_MAX / 2
    R = COORD_MAX * (0.25 + 0.2 * float(s.unit(1)[0]))
    m_out = 16 + int(s.below(1, 497)[0])
    m_out = max(3, min(m_out, n // 2))
    segs = _star_polygon(s, cx, cy, R, m_out)
    nh = int(s.below(1, 5)[0])
    holes = []
    budget = n // 2 - m_out
    for _ in range(nh):
        mh = 16 + int(s.below(1, 49)[0])
        if mh > budget:
            break
        for _try in range(20):
            rh = R * (0.05 + 0.1 * float(s.unit(1)[0]))
            ang, dist = s.unit(2).tolist()
            d = (0.6 * R - rh) * dist
            hx = cx + d * math.cos(2 * math.pi * ang)
            hy = cy + d * math.sin(2 * math.pi * ang)
            if all(math.hypot(hx - ox, hy - oy) > rh + orr + 2 for ox, oy, orr in holes):
                holes.append((hx, hy, rh))
                segs += _star_polygon(s, hx, hy, rh, mh)
                budget -= mh
                break
    ninner = max(0, n - len(segs))
    # uniform in the disk of radius 0.65 R, outside the holes' circles
    pts = []
    have = 0
    while have < ninner:
        want = (ninner - have) * 2 + 16
        u = s.unit(2 * want).reshape(-1, 2)
        px = cx + (u[:, 0] * 2 - 1) * 0.65 * R
        py = cy + (u[:, 1] * 2 - 1) * 0.65 * R
        ok = (px - cx) ** 2 + (py - cy) ** 2 < (0.65 * R) ** 2
        for hx, hy, rh in holes:
            ok &= (px - hx) ** 2 + (py - hy) ** 2 > (rh * 1.05) ** 2
        sel = np.stack([np.floor(px[ok]), np.floor(py[ok])], axis=1).astype(np.int64)[: ninner - have]
        pts.append(sel)
        have += len(sel)
    if not pts:
        pts = [np.zeros((0, 2), dtype=np.int64)]
    return _finish(pts, segs)


def batched_instances(count: int = 4096, nmin: int = 64, nmax: int = 65536, seed: int = 6):
    """`count` independent star instances, point counts log-uniform in [nmin, nmax]."""
    s = Stream(seed)
    out = []
    lo, hi = math.log(nmin), math.log(nmax)
    for _ in range(count):
        n = int(round(math.exp(lo + (hi - lo) * float(s.unit(1)[0]))))
        out.append(star_instance(s, n))
    return out


def polygon_instance(s: Stream, n: int, frac_lo: float = 0.01, frac_hi: float = 0.05):
    """Uniform points plus random disjoint polygon boundaries whose edges are
    1-5% of the point count."""
    pts = np.stack([s.bits(n, COORD_BITS), s.bits(n, COORD_BITS)], axis=1)
    want = max(3, int(n * (frac_lo + (frac_hi - frac_lo) * float(s.unit(1)[0]))))
    segs = []
    boxes = []
    while len(segs) < want:
        k = int(min(want - len(segs), 16 + int(s.below(1, 85)[0])))
        k = max(k, 3)
        before = len(boxes)
        segs += _circle_segments(s, before + 1, k, 64, 8192, taken=boxes)
        if len(boxes) == before:
            break
    return _finish([pts], segs)


def serving_requests(count: int = 256, qps: float = 50.0, nmin: float = 1e3, nmax: float = 1e6, seed: int = 7):
    """Poisson arrivals: list of (arrival_s, points, constraints)."""
    s = Stream(seed)
    t = 0.0
    lo, hi = math.log(nmin), math.log(nmax)
    out = []
    for _ in range(count):
        t += -math.log(1.0 - float(s.unit(1)[0])) / qps
        n = int(round(math.exp(lo + (hi - lo) * float(s.unit(1)[0]))))
        pts, cons = polygon_instance(s, n)
        out.append((t, pts, cons))
    return out
