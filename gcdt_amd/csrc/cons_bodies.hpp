// Per-item bodies of the constraint-recovery kernels (cons.hip).
#pragma once
#include "mesh_bodies.hpp"

namespace gcdt {

struct Cons {
    Mesh m;
    int* vt;                       // vertex -> an incident triangle
    unsigned long long* owner;     // per triangle: round key of the owning piece
    int* stamp;                    // per triangle: epoch of its last rewrite
    int* cavof;                    // per triangle: winner index of that rewrite
    int* err;
    int steiner = 0;               // crossing pieces are deferred instead of failing
    int* defer = nullptr;          // per piece: 1 when deferred (steiner mode)
};

__host__ __device__ inline bool on_ray(int2 A, int2 U, int2 B)
{
    // U collinear with A->B (caller checked) and in the direction of B
    return (long long)(U.x - A.x) * (B.x - A.x) + (long long)(U.y - A.y) * (B.y - A.y) > 0;
}

// Rotate around vertex `cur` to the triangle whose corner at `cur` contains
// the direction cur->b.  Returns 0 with (t,s) when the segment leaves through
// the opposite edge, 1 when the vertex u = tri[t][s+1] lies on the ray (the
// segment runs along edge (cur,u)), -1 on failure.
__host__ __device__ inline int find_wedge(const Cons& c, int cur, int b, int& t, int& s)
{
    const Mesh& m = c.m;
    int2 P = m.pts[cur], B = m.pts[b];
    int t0 = c.vt[cur];
    if (t0 < 0) return -1;
    t = t0;
    for (int guard = 0; guard < 1 << 20; guard++) {
        int3 tv = m.tri[t];
        s = slot_of(tv, cur);
        if (s < 0) return -1;
        int u = get3(tv, nx3(s)), w = get3(tv, pv3(s));
        int2 U = m.pts[u];
        int ou = orient2d(P, U, B);
        if (ou == 0 && on_ray(P, U, B)) return 1;
        int ow = orient2d(P, m.pts[w], B);
        if (ou > 0 && ow < 0) return 0;
        int n = A3(m, t, nx3(s));  // across (cur, w): counter-clockwise around cur
        if (n < 0) return -1;
        t = n;
        if (t == t0) return -1;
    }
    return -1;
}

__host__ __device__ inline void mark_pair(const Cons& c, int t, int k)
{
    const Mesh& m = c.m;
    at_or(&m.cflag[t], 1 << k);
    int n = A3(m, t, k);
    if (n >= 0) {
        int3 tv = m.tri[t];
        int sn = slot_other(m.tri[n], get3(tv, nx3(k)), get3(tv, pv3(k)));
        at_or(&m.cflag[n], 1 << sn);
    }
}

// A walk along a segment through the triangles it crosses.  It holds the
// current triangle's vertices, neighbours and constraint bits, so crossing an
// edge costs two dependent loads: the neighbour's vertices, then its
// neighbours, its constraint bits and the apex's point together.
struct Cross {
    int t = -1;
    int3 tv, ta;
    int cf = 0;
    __host__ __device__ void start(const Mesh& m, int t0)
    {
        t = t0;
        tv = m.tri[t0];
        ta = m.adj[t0];
        cf = m.cflag[t0];
    }
    // Crosses the edge (r,l) of t: n is the neighbour, q its apex and P the
    // apex's point; the walk moves to n when n >= 0.  Returns the constraint
    // bit of the crossed edge.
    __host__ __device__ int step(const Mesh& m, int r, int l, int& n, int& q, int2& P)
    {
        const int k = slot_other(tv, r, l);
        const int bit = (cf >> k) & 1;
        n = get3(ta, k);
        if (n >= 0) {
            const int3 nv = m.tri[n];
            q = get3(nv, slot_other(nv, r, l));
            const int3 na = m.adj[n];
            const int ncf = m.cflag[n];
            P = m.pts[q];
            t = n;
            tv = nv;
            ta = na;
            cf = ncf;
        }
        return bit;
    }
};

// ---------------------------------------------------------------- split pass

// Walk segment i; emit its pieces (when out != nullptr) at out[base..] and
// return the piece count.  Pieces that are existing edges are marked now and
// written with a negative first vertex (-(a+1)) so the rounds skip them.
__host__ __device__ inline int split_walk(const Cons& c, int a, int b, int2* out)
{
    const Mesh& m = c.m;
    int2 PA = m.pts[a], PB = m.pts[b];
    int cur = a, np = 0;
    for (int guard = 0; guard < (1 << 26) && cur != b; guard++) {
        int t, s;
        int r = find_wedge(c, cur, b, t, s);
        if (r < 0) {
            at_or(c.err, MERR_WALK);
            return np;
        }
        int3 tv = m.tri[t];
        if (r == 1) {
            int u = get3(tv, nx3(s));
            if (out) {
                mark_pair(c, t, pv3(s));
                out[np] = make_int2(-(cur + 1), u);
            }
            np++;
            cur = u;
            continue;
        }
        int rr = get3(tv, nx3(s)), ll = get3(tv, pv3(s));
        Cross wk;
        wk.start(m, t);
        int next = -1;
        for (int g2 = 0; g2 < (1 << 26); g2++) {
            int n, q;
            int2 PQ;
            // a constrained edge crossed: an error unless crossings are
            // resolved later (the split only looks for vertices on the segment)
            if (wk.step(m, rr, ll, n, q, PQ) && !c.steiner) {
                at_or(c.err, MERR_CROSSING);
                return np;
            }
            if (n < 0) {
                at_or(c.err, MERR_WALK);
                return np;
            }
            if (q == b) {
                next = b;
                break;
            }
            int o = orient2d(PA, PB, PQ);
            if (o == 0) {
                next = q;
                break;
            }
            if (o > 0) ll = q;
            else rr = q;
        }
        if (next < 0) {
            at_or(c.err, MERR_WALK);
            return np;
        }
        if (out) out[np] = make_int2(cur, next);
        np++;
        cur = next;
    }
    return np;
}

__host__ __device__ inline unsigned long long cav_key(unsigned epoch, int piece)
{
    return ((unsigned long long)(~epoch) << 32) | (unsigned)piece;
}

// Walk piece (a,b) from its wedge at a; calls f(tri) for each crossed
// triangle.  Returns the crossed count, 0 if (a,b) already is an edge (marked
// here when mark is set), -1 on error, -2 when it meets a constrained edge
// and crossings are deferred (c.steiner).
template <class F>
__host__ __device__ inline int cav_walk(const Cons& c, int a, int b, bool mark, int* t0out, F&& f)
{
    const Mesh& m = c.m;
    int2 PA = m.pts[a], PB = m.pts[b];
    int t, s;
    int r = find_wedge(c, a, b, t, s);
    if (r < 0) return -1;
    int3 tv = m.tri[t];
    if (r == 1) {
        if (get3(tv, nx3(s)) != b) return -1;  // a vertex inside the piece: not split
        if (mark) mark_pair(c, t, pv3(s));
        return 0;
    }
    if (t0out) *t0out = t;
    int rr = get3(tv, nx3(s)), ll = get3(tv, pv3(s));
    int cnt = 1;
    f(t);
    Cross wk;
    wk.start(m, t);
    for (int g = 0; g < (1 << 26); g++) {
        int n, q;
        int2 PQ;
        if (wk.step(m, rr, ll, n, q, PQ)) {
            if (c.steiner) return -2;
            at_or(c.err, MERR_CROSSING);
            return -1;
        }
        if (n < 0) return -1;
        f(n);
        cnt++;
        if (q == b) return cnt;
        int o = orient2d(PA, PB, PQ);
        if (o == 0) return -1;
        if (o > 0) ll = q;
        else rr = q;
    }
    return -1;
}

// R2: winners reserve scratch, losers stay queued
__host__ __device__ inline long long cav_scratch_ints(int K) { return 13LL * K + 20; }

struct Win {
    int piece;
    int t0;
    int K;
    int off;  // scratch offset (ints)
};

// The Delaunay vertex of base (A, B) among the polygon vertices lo..hi-1
// (coordinates in PX/PY): the one whose circle with A and B holds no other
// strictly inside, the first in order among cocircular ones.  One thread
// scans them in order, replacing its candidate by any vertex strictly inside
// the candidate's circle.
struct DvScan {
    __host__ __device__ static int find(const int* PX, const int* PY, int lo, int hi, int2 PA, int2 PB)
    {
        int ci = lo;
        int2 PC = make_int2(PX[lo], PY[lo]);
        for (int i = lo + 1; i < hi; i++) {
            const int2 PI = make_int2(PX[i], PY[i]);
            if (incircle(PA, PB, PC, PI) > 0) {
                ci = i;
                PC = PI;
            }
        }
        return ci;
    }
};

// R3 body: collect winner wi's cavity and retriangulate it in place.
// Returns false (and flags MERR_CAVITY) on an inconsistent cavity.  DV finds
// each Delaunay vertex (DvScan, or a wave-wide search whose lanes all run this
// body for the same cavity).
template <class DV = DvScan>
__host__ __device__ inline bool cav_retri_one(const Cons& c, const int2* pieces, const Win* wins, int wi, int* scr,
                                              unsigned epoch)
{
    const Mesh& m = c.m;
        Win w = wins[wi];
        int K = w.K;
        int* slots = scr + w.off;         // K
        int* Lv = slots + K;              // K (left chain, walk order)
        int* Rv = Lv + K;                 // K (right chain, walk order)
        int* Lb = Rv + K;                 // K+1 outer ids, left edges
        int* Lc = Lb + (K + 1);           // K+1 cflags
        int* Rb = Lc + (K + 1);           // K+1 outer ids, right edges (walk order)
        int* Rc = Rb + (K + 1);           // K+1
        int* stk = Rc + (K + 1);          // 4 * (K+1)
        int* PX = stk + 4 * (K + 1);      // K+2: one pseudo-polygon's vertex coordinates
        int* PY = PX + (K + 2);           // K+2
        int2 pc = pieces[w.piece];
        int a = pc.x, b = pc.y;
        int2 PA = m.pts[a], PB = m.pts[b];
        // ---- collect (walk from the owned start triangle)
        int t = w.t0;
        int3 tv = m.tri[t];
        int s = slot_of(tv, a);
        if (s < 0) {
            at_or(c.err, MERR_CAVITY);
            return false;
        }
        int rr = get3(tv, nx3(s)), ll = get3(tv, pv3(s));
        int ft = m.cflag[t];
        int nl = 0, nr = 0, nlb = 0, nrb = 0, ns = 0;
        slots[ns++] = t;
        Lv[nl++] = ll;
        Rv[nr++] = rr;
        Lb[nlb] = A3(m, t, nx3(s));  // (w,a) opposite u
        Lc[nlb++] = (ft >> nx3(s)) & 1;
        Rb[nrb] = A3(m, t, pv3(s));  // (a,u) opposite w
        Rc[nrb++] = (ft >> pv3(s)) & 1;
        bool bad = false;
        Cross wk;
        wk.start(m, t);
        for (;;) {
            int n, q;
            int2 PQ;
            wk.step(m, rr, ll, n, q, PQ);
            if (n < 0 || ns >= K) {
                bad = true;
                break;
            }
            slots[ns++] = n;
            // the walk now holds n's vertices, neighbours and constraint bits
            const int3 nv = wk.tv, na = wk.ta;
            const int fn = wk.cf;
            int kr = slot_of(nv, rr), kl = slot_of(nv, ll);
            if (kr < 0 || kl < 0) {
                bad = true;
                break;
            }
            if (q == b) {
                Lb[nlb] = get3(na, kr);  // (ll,b) opposite rr
                Lc[nlb++] = (fn >> kr) & 1;
                Rb[nrb] = get3(na, kl);  // (rr,b) opposite ll
                Rc[nrb++] = (fn >> kl) & 1;
                break;
            }
            int o = orient2d(PA, PB, PQ);
            if (o > 0) {  // boundary (ll,q) opposite rr
                Lb[nlb] = get3(na, kr);
                Lc[nlb++] = (fn >> kr) & 1;
                Lv[nl++] = q;
                ll = q;
            } else if (o < 0) {  // boundary (rr,q) opposite ll
                Rb[nrb] = get3(na, kl);
                Rc[nrb++] = (fn >> kl) & 1;
                Rv[nr++] = q;
                rr = q;
            } else {
                bad = true;
                break;
            }
        }
        if (bad || ns != K || nl + nr != K) {
            at_or(c.err, MERR_CAVITY);
            return false;
        }
        // ---- triangulate both pseudo-polygons
        // left polygon  poly[0]=a, poly[1..nl]=Lv, poly[nl+1]=b; edge i = Lb[i]
        // right polygon poly[0]=b, poly[1..nr]=Rv reversed, poly[nr+1]=a;
        //               edge i = Rb[nr - i]
        int e = 0;
        int topL = -1, topR = -1;
        for (int side = 0; side < 2; side++) {
            int M = side == 0 ? nl : nr;
            // the polygon's coordinates in order, so the Delaunay-vertex scans
            // read consecutive words instead of a vertex id and then its point
            for (int i = 0; i <= M + 1; i++) {
                const int v = side == 0 ? (i == 0 ? a : (i == M + 1 ? b : Lv[i - 1]))
                                        : (i == 0 ? b : (i == M + 1 ? a : Rv[M - i]));
                const int2 P = m.pts[v];
                PX[i] = P.x;
                PY[i] = P.y;
            }
            int sp = 0;
            stk[0] = 1;
            stk[1] = M + 1;
            stk[2] = -1;
            stk[3] = 0;
            sp = 1;
            while (sp > 0) {
                sp--;
                int lo = stk[4 * sp], hi = stk[4 * sp + 1], par = stk[4 * sp + 2], pslot = stk[4 * sp + 3];
                auto poly = [&](int i) -> int {
                    if (side == 0) return i == 0 ? a : (i == M + 1 ? b : Lv[i - 1]);
                    return i == 0 ? b : (i == M + 1 ? a : Rv[M - i]);
                };
                auto bnd_t = [&](int i) -> int { return side == 0 ? Lb[i] : Rb[M - i]; };
                auto bnd_c = [&](int i) -> int { return side == 0 ? Lc[i] : Rc[M - i]; };
                int A = poly(lo - 1), B = poly(hi);
                const int ci = DV::find(PX, PY, lo, hi, make_int2(PX[lo - 1], PY[lo - 1]), make_int2(PX[hi], PY[hi]));
                int C = poly(ci);
                int T = slots[e++];
                int3 nt = make_int3(A, B, C);
                int3 na = make_int3(-1, -1, par);
                int cf = 0;
                if (par >= 0) A3(m, par, pslot) = T;
                else if (side == 0) topL = T;
                else topR = T;
                if (ci > lo) {
                    stk[4 * sp] = lo;
                    stk[4 * sp + 1] = ci;
                    stk[4 * sp + 2] = T;
                    stk[4 * sp + 3] = 1;
                    sp++;
                } else {
                    na.y = adj_mark(bnd_t(lo - 1));
                    cf |= bnd_c(lo - 1) << 1;
                }
                if (ci + 1 < hi) {
                    stk[4 * sp] = ci + 1;
                    stk[4 * sp + 1] = hi;
                    stk[4 * sp + 2] = T;
                    stk[4 * sp + 3] = 0;
                    sp++;
                } else {
                    na.x = adj_mark(bnd_t(ci));
                    cf |= bnd_c(ci);
                }
                m.tri[T] = nt;
                m.adj[T] = na;
                m.cflag[T] = cf;
                c.stamp[T] = (int)epoch;
                c.cavof[T] = (int)wi;
                c.vt[A] = T;
                c.vt[B] = T;
                c.vt[C] = T;
            }
        }
        // the constraint edge joins the two top triangles (slot 2 = edge A,B)
        A3(m, topL, 2) = topR;
        A3(m, topR, 2) = topL;
        m.cflag[topL] |= 4;
        m.cflag[topR] |= 4;
    return true;
}

// R4 body: resolve the marked outer adjacency of rewritten triangle t.
__host__ __device__ inline void cav_fix_one(const Cons& c, const Win* wins, const int* scr, int t, unsigned epoch)
{
    const Mesh& m = c.m;
        int3 tv = m.tri[t];
        int3 ta = m.adj[t];
        for (int k = 0; k < 3; k++) {
            int av = get3(ta, k);
            if (!adj_is_mark(av)) continue;
            int nb = adj_unmark(av);
            int u = get3(tv, nx3(k)), v = get3(tv, pv3(k));
            if (c.stamp[nb] == (int)epoch) {
                Win w = wins[c.cavof[nb]];
                const int* sl = scr + w.off;
                int found = -1;
                for (int j = 0; j < w.K; j++)
                    if (sl[j] != t && has2(m.tri[sl[j]], u, v)) {
                        found = sl[j];
                        break;
                    }
                if (found < 0) at_or(c.err, MERR_ADJ);
                set3(ta, k, found);
            } else {
                set3(ta, k, nb);
                A3(m, nb, slot_other(m.tri[nb], u, v)) = t;
            }
        }
        m.adj[t] = ta;
}

}  // namespace gcdt
