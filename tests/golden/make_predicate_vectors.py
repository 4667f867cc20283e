"""Writes tests/golden/predicates.json: orient2d / inCircle sign vectors
computed with Python big integers from the closed-form determinants.

orient2d(a,b,c) = sign((a-c) x (b-c))      (include/math_base.cuh:68-76)
inCircle(a,b,c,p) = sign(det | a-p  |a-p|^2 |
                             | b-p  |b-p|^2 |
                             | c-p  |c-p|^2 |)   (include/math_base.cuh:267-328)

Cases: random at 2^20 and 2^30 magnitudes, the extremes {0, 2^30-1}, exact
collinear triples, exact cocircular quadruples (lattice rectangles and
Pythagorean-triple circles), and near-degenerate perturbations (+-1) of both.
Run: python tests/golden/make_predicate_vectors.py
"""
import json
import os
import random

R = (1 << 30) - 1


def orient(a, b, c):
    v = (a[0] - c[0]) * (b[1] - c[1]) - (a[1] - c[1]) * (b[0] - c[0])
    return (v > 0) - (v < 0)


def incircle(a, b, c, d):
    m = []
    for q in (a, b, c):
        x, y = q[0] - d[0], q[1] - d[1]
        m.append((x, y, x * x + y * y))
    det = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
           - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
           + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))
    return (det > 0) - (det < 0)


def clamp(v):
    return max(0, min(R, v))


def main():
    rnd = random.Random(20240601)
    ori, inc = [], []

    def P(bits):
        return (rnd.randrange(1 << bits), rnd.randrange(1 << bits))

    for bits in (4, 20, 30):
        for _ in range(300):
            a, b, c, d = P(bits), P(bits), P(bits), P(bits)
            ori.append([a, b, c])
            inc.append([a, b, c, d])
    ext = [0, 1, R - 1, R]
    for _ in range(300):
        pts = [(rnd.choice(ext), rnd.choice(ext)) for _ in range(4)]
        ori.append(pts[:3])
        inc.append(pts)
    # collinear triples and +-1 perturbations
    for _ in range(300):
        a = (rnd.randrange(R), rnd.randrange(R))
        dx, dy = rnd.randrange(-2000, 2000), rnd.randrange(-2000, 2000)
        k1, k2 = rnd.randrange(-20000, 20000), rnd.randrange(-20000, 20000)
        b = (clamp(a[0] + k1 * dx), clamp(a[1] + k1 * dy))
        c = (clamp(a[0] + k2 * dx), clamp(a[1] + k2 * dy))
        ori.append([a, b, c])
        ori.append([a, b, (clamp(c[0] + 1), c[1])])
        ori.append([a, b, (c[0], clamp(c[1] - 1))])
    # cocircular: Pythagorean circles and lattice rectangles
    triples = [(3, 4, 5), (5, 12, 13), (8, 15, 17), (7, 24, 25), (20, 21, 29), (9, 40, 41)]
    for _ in range(300):
        p, q, r = rnd.choice(triples)
        s = rnd.randrange(1, (1 << 24) // r)
        cx, cy = rnd.randrange(r * s, R - r * s), rnd.randrange(r * s, R - r * s)
        circ = [(p, q), (q, -p), (-p, -q), (-q, p), (r, 0), (0, r), (-r, 0), (0, -r), (p, -q), (-p, q)]
        pick = rnd.sample(circ, 4)
        pts = [(cx + s * x, cy + s * y) for x, y in pick]
        a, b, c = pts[0], pts[1], pts[2]
        if orient(a, b, c) < 0:
            a, b = b, a
        if orient(a, b, c) == 0:
            continue
        inc.append([a, b, c, pts[3]])
        inc.append([a, b, c, (pts[3][0] + 1, pts[3][1])])
        inc.append([a, b, c, (pts[3][0], pts[3][1] - 1)])
    for _ in range(300):
        x0, y0 = rnd.randrange(R // 2), rnd.randrange(R // 2)
        w, h = rnd.randrange(1, R // 2), rnd.randrange(1, R // 2)
        a, b, c, d = (x0, y0), (x0 + w, y0), (x0 + w, y0 + h), (x0, y0 + h)
        inc.append([a, b, c, d])
        inc.append([a, b, c, (d[0] + 1, d[1])])
        inc.append([a, b, c, (d[0] - 1, d[1])])
    data = {
        "orient2d": [{"p": t, "sign": orient(*t)} for t in ori],
        "incircle": [{"p": t, "sign": incircle(*t)} for t in inc],
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "predicates.json")
    with open(out, "w") as f:
        json.dump(data, f, separators=(",", ":"))
    print(out, len(data["orient2d"]), len(data["incircle"]))


if __name__ == "__main__":
    main()
