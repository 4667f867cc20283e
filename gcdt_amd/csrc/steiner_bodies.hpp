// Crossing constraints resolved with exact Steiner points (one device thread).
//
// The reference has no answer for two constraint segments that cross: its
// walk meets a constrained edge and the piece is dropped
// (addConstraints.cu:768-774,942-946).  Here the cavity rounds retire such a
// piece instead ("deferred"), the mesh is finished as the CDT of the other
// pieces, and this pass inserts the deferred pieces one after another into
// that CDT:
//   * walk the piece (a, b) from a towards b; each constrained edge it meets
//     is split at the exact crossing point s = line(a,b) x line(support), a
//     new vertex with homogeneous coordinates (exact.hpp), and the four
//     edges around s are made locally Delaunay again by Lawson flips that
//     never flip a constrained edge (incremental CDT insertion of a vertex
//     on a constrained edge);
//   * the stretch from the previous vertex on the piece to s crosses only
//     unconstrained edges and holds no vertex: it is recovered by flipping
//     the edges that cross it (a queue of crossing edges: flip an edge whose
//     quadrilateral is strictly convex, keep the new edge in the queue while
//     it still crosses), marked constrained, and the new edges are made
//     locally Delaunay by Lawson flips;
//   * a vertex met exactly on the piece (a crossing of two other pieces on
//     which this one passes too) ends a stretch the same way.
// A triangulation whose unconstrained edges are all locally Delaunay is the
// CDT of its vertices and constrained edges, so the result is the CDT of the
// input points, the corners and the crossing points, with every piece a chain
// of constrained edges through its crossings.
//
// Support of a constrained edge: the original piece (two grid vertices) it
// lies on.  Every Steiner vertex records the pieces through it (up to
// STN_SUPS), so a crossing point is always computed from two pieces with grid
// endpoints and its coordinates stay below 2^92 (W below 2^61).
#pragma once
#include "cons_bodies.hpp"
#include "exact.hpp"

namespace gcdt {

constexpr int STN_SUPS = 4;
enum { STN_T = 0, STN_S = 1, STN_NCTR = 2 };

struct Stn {
    int3* tri;
    int3* adj;
    int* cflag;
    const int2* pts;  // grid vertices, ids below nv0
    int nv0;
    HPt* hp;          // Steiner vertex nv0 + k at hp[k]
    int2* sup;        // STN_SUPS pieces (grid vertex ids) per Steiner vertex; x < 0 when unused
    int* vt;          // vertex -> an incident triangle
    int* scr;         // scratch ints
    long long scr_cap;
    int* ctr;         // STN_T triangles in use, STN_S Steiner vertices
    int tcap, scap;
    int* err;
};

__host__ __device__ inline HPt stn_pt(const Stn& s, int v) { return v < s.nv0 ? hpt_grid(s.pts[v]) : s.hp[v - s.nv0]; }

__host__ __device__ inline int stn_orient(const Stn& s, int a, int b, int c)
{
    if (a < s.nv0 && b < s.nv0 && c < s.nv0) return orient2d(s.pts[a], s.pts[b], s.pts[c]);
    return orient_h(stn_pt(s, a), stn_pt(s, b), stn_pt(s, c));
}

__host__ __device__ inline int stn_incircle(const Stn& s, int a, int b, int c, int d)
{
    if (a < s.nv0 && b < s.nv0 && c < s.nv0 && d < s.nv0) return incircle(s.pts[a], s.pts[b], s.pts[c], s.pts[d]);
    return incircle_h(stn_pt(s, a), stn_pt(s, b), stn_pt(s, c), stn_pt(s, d));
}

// u collinear with p -> b: true when u lies on the ray from p through b
__host__ __device__ inline bool stn_on_ray(const Stn& s, int p, int u, int b)
{
    HPt P = stn_pt(s, p), U = stn_pt(s, u), B = stn_pt(s, b);
    int dx = hpt_cmp_x(B, P);
    if (dx != 0) return hpt_cmp_x(U, P) == dx;
    return hpt_cmp_y(U, P) == hpt_cmp_y(B, P);
}

// The triangle t and slot k (opposite) of edge (u, v); false if no such edge.
// Rotates around u both ways (u may be a corner on the hull).
__host__ __device__ inline bool stn_find_edge(const Stn& s, int u, int v, int& t, int& k)
{
    const int t0 = s.vt[u];
    if (t0 < 0) return false;
    for (int dir = 0; dir < 2; dir++) {
        int x = t0;
        for (int g = 0; g < (1 << 16); g++) {
            const int3 tv = s.tri[x];
            const int su = slot_of(tv, u);
            if (su < 0) {
                at_or(s.err, MERR_ADJ);
                return false;
            }
            if (slot_of(tv, v) >= 0) {
                t = x;
                k = slot_other(tv, u, v);
                return true;
            }
            const int n = get3(s.adj[x], dir == 0 ? nx3(su) : pv3(su));
            if (n < 0) break;
            x = n;
            if (x == t0) return false;
        }
    }
    return false;
}

// find_wedge (cons_bodies.hpp) over homogeneous vertices: the triangle t
// whose corner at cur (slot k) holds the direction cur -> b.  0: the segment
// leaves through the opposite edge; 1: the next vertex counter-clockwise lies
// on the ray; -1: failure.  cur is never on the hull (constraint endpoints
// and crossings lie inside the square).
__host__ __device__ inline int stn_find_wedge(const Stn& s, int cur, int b, int& t, int& k)
{
    const int t0 = s.vt[cur];
    if (t0 < 0) return -1;
    t = t0;
    for (int g = 0; g < (1 << 16); g++) {
        const int3 tv = s.tri[t];
        k = slot_of(tv, cur);
        if (k < 0) return -1;
        const int u = get3(tv, nx3(k)), w = get3(tv, pv3(k));
        const int ou = stn_orient(s, cur, u, b);
        if (ou == 0 && stn_on_ray(s, cur, u, b)) return 1;
        const int ow = stn_orient(s, cur, w, b);
        if (ou > 0 && ow < 0) return 0;
        const int n = get3(s.adj[t], nx3(k));
        if (n < 0) return -1;
        t = n;
        if (t == t0) return -1;
    }
    return -1;
}

// constrained bit of edge k of t, and of the same edge in the neighbour
__host__ __device__ inline void stn_set_c(const Stn& s, int t, int k)
{
    s.cflag[t] |= 1 << k;
    const int n = get3(s.adj[t], k);
    if (n >= 0) {
        const int3 tv = s.tri[t];
        s.cflag[n] |= 1 << slot_other(s.tri[n], get3(tv, nx3(k)), get3(tv, pv3(k)));
    }
}

__host__ __device__ inline void stn_repoint(const Stn& s, int nb, int u, int v, int to)
{
    if (nb >= 0) ((int*)s.adj)[3 * (long long)nb + slot_other(s.tri[nb], u, v)] = to;
}

// Flips the edge opposite slot k of t:
// t = (p,u,v), n = (o,v,u)  ->  t = (p,u,o), n = (p,o,v).
__host__ __device__ inline void stn_flip(const Stn& s, int t, int k)
{
    const int kn = nx3(k), kp = pv3(k);
    const int3 tv = s.tri[t], ta = s.adj[t];
    const int n = get3(ta, k);
    const int3 nv = s.tri[n], na = s.adj[n];
    const int p = get3(tv, k), u = get3(tv, kn), v = get3(tv, kp);
    const int j = slot_other(nv, u, v);
    const int jn = nx3(j), jp = pv3(j);
    const int o = get3(nv, j);
    const int cft = s.cflag[t], cfn = s.cflag[n];
    const int uo = get3(na, jn), ov = get3(na, jp), pu = get3(ta, kp), vp = get3(ta, kn);
    s.tri[t] = make_int3(p, u, o);
    s.adj[t] = make_int3(uo, n, pu);
    s.cflag[t] = ((cfn >> jn) & 1) | (((cft >> kp) & 1) << 2);
    s.tri[n] = make_int3(p, o, v);
    s.adj[n] = make_int3(ov, vp, t);
    s.cflag[n] = ((cfn >> jp) & 1) | (((cft >> kn) & 1) << 1);
    stn_repoint(s, uo, u, o, t);
    stn_repoint(s, vp, v, p, n);
    s.vt[p] = t;
    s.vt[u] = t;
    s.vt[o] = t;
    s.vt[v] = n;
}

// Splits the edge opposite slot k of t at the new vertex sv (strictly inside
// the edge): t = (p,u,v), n = (q,v,u) become t = (p,u,sv), t3 = (p,sv,v),
// n = (q,v,sv), t4 = (q,sv,u); both halves keep the edge's constrained bit.
// The four outer edges are pushed on stk (vertex pairs) at *sp.
__host__ __device__ inline bool stn_split(const Stn& s, int t, int k, int sv, int* stk, long long& sp)
{
    const int kn = nx3(k), kp = pv3(k);
    const int3 tv = s.tri[t], ta = s.adj[t];
    const int n = get3(ta, k);
    if (n < 0) {
        at_or(s.err, MERR_WALK);
        return false;
    }
    const int T = s.ctr[STN_T];
    if (T + 2 > s.tcap) {
        at_or(s.err, MERR_CAPACITY);
        return false;
    }
    s.ctr[STN_T] = T + 2;
    const int t3 = T, t4 = T + 1;
    const int3 nv = s.tri[n], na = s.adj[n];
    const int p = get3(tv, k), u = get3(tv, kn), v = get3(tv, kp);
    const int j = slot_other(nv, u, v);
    const int q = get3(nv, j);
    const int cft = s.cflag[t], cfn = s.cflag[n];
    const int c = (cft >> k) & 1;
    const int pu = get3(ta, kp), vp = get3(ta, kn);
    const int qv = get3(na, pv3(j)), uq = get3(na, nx3(j));  // n = (q,v,u): opposite u, opposite v
    s.tri[t] = make_int3(p, u, sv);
    s.adj[t] = make_int3(t4, t3, pu);
    s.cflag[t] = c | (((cft >> kp) & 1) << 2);
    s.tri[t3] = make_int3(p, sv, v);
    s.adj[t3] = make_int3(n, vp, t);
    s.cflag[t3] = c | (((cft >> kn) & 1) << 1);
    s.tri[n] = make_int3(q, v, sv);
    s.adj[n] = make_int3(t3, t4, qv);
    s.cflag[n] = c | (((cfn >> pv3(j)) & 1) << 2);
    s.tri[t4] = make_int3(q, sv, u);
    s.adj[t4] = make_int3(t, uq, n);
    s.cflag[t4] = c | (((cfn >> nx3(j)) & 1) << 1);
    stn_repoint(s, vp, v, p, t3);
    stn_repoint(s, uq, u, q, t4);
    s.vt[sv] = t;
    s.vt[p] = t;
    s.vt[u] = t;
    s.vt[v] = t3;
    s.vt[q] = n;
    if (2 * (sp + 4) > s.scr_cap) {
        at_or(s.err, MERR_SCRATCH);
        return false;
    }
    const int e[8] = {p, u, v, p, q, v, u, q};
    for (int i = 0; i < 8; i++) stk[2 * sp + i] = e[i];
    sp += 4;
    return true;
}

// Lawson flips from the edges on the stack until every edge popped is
// locally Delaunay, constrained or on the hull.
__host__ __device__ inline bool stn_legalize(const Stn& s, int* stk, long long cap_pairs, long long sp)
{
    long long guard = 0;
    while (sp > 0) {
        if (++guard > (1LL << 28)) {
            at_or(s.err, MERR_CAVITY);
            return false;
        }
        sp--;
        const int a = stk[2 * sp], b = stk[2 * sp + 1];
        int t, k;
        if (!stn_find_edge(s, a, b, t, k)) continue;  // flipped away meanwhile
        if ((s.cflag[t] >> k) & 1) continue;
        const int n = get3(s.adj[t], k);
        if (n < 0) continue;
        const int3 tv = s.tri[t], nv = s.tri[n];
        const int p = get3(tv, k), u = get3(tv, nx3(k)), v = get3(tv, pv3(k));
        const int o = get3(nv, slot_other(nv, u, v));
        if (stn_incircle(s, tv.x, tv.y, tv.z, o) <= 0) continue;
        stn_flip(s, t, k);
        if (sp + 4 > cap_pairs) {
            at_or(s.err, MERR_SCRATCH);
            return false;
        }
        const int e[8] = {p, u, u, o, o, v, v, p};
        for (int i = 0; i < 8; i++) stk[2 * sp + i] = e[i];
        sp += 4;
    }
    return true;
}

// Makes (x, y) a constrained edge.  The open segment holds no vertex and
// crosses no constrained edge (both checked).
__host__ __device__ inline bool stn_recover(const Stn& s, int x, int y)
{
    int t, k;
    if (stn_find_edge(s, x, y, t, k)) {
        stn_set_c(s, t, k);
        return true;
    }
    // scratch: the queue of crossing edges (circular), then the new edges
    const long long half = s.scr_cap / 2;
    int* Q = s.scr;
    int* L = s.scr + half;
    const long long qcap = half / 2, lcap = half / 2;
    long long nq = 0;
    int w;
    if (stn_find_wedge(s, x, y, t, w) != 0) {
        at_or(s.err, MERR_WALK);
        return false;
    }
    int rr = get3(s.tri[t], nx3(w)), ll = get3(s.tri[t], pv3(w));
    for (long long g = 0;; g++) {
        if (nq >= qcap || g > (1LL << 26)) {
            at_or(s.err, MERR_SCRATCH);
            return false;
        }
        Q[2 * nq] = rr;
        Q[2 * nq + 1] = ll;
        nq++;
        const int kk = slot_other(s.tri[t], rr, ll);
        const int n = get3(s.adj[t], kk);
        if (n < 0 || ((s.cflag[t] >> kk) & 1)) {
            at_or(s.err, MERR_WALK);
            return false;
        }
        const int3 nv = s.tri[n];
        const int qv = get3(nv, slot_other(nv, rr, ll));
        if (qv == y) break;
        const int o = stn_orient(s, x, y, qv);
        if (o == 0) {
            at_or(s.err, MERR_WALK);
            return false;
        }
        if (o > 0) ll = qv;
        else rr = qv;
        t = n;
    }
    // flip crossing edges until none is left
    long long head = 0, cnt = nq, nl = 0;
    const long long limit = 64 * (nq + 1) * (nq + 1) + 4096;
    for (long long g = 0; cnt > 0; g++) {
        if (g > limit) {
            at_or(s.err, MERR_CAVITY);
            return false;
        }
        const int a = Q[2 * head], b = Q[2 * head + 1];
        head = (head + 1) % qcap;
        cnt--;
        if (!stn_find_edge(s, a, b, t, k)) {
            at_or(s.err, MERR_CAVITY);
            return false;
        }
        const int n = get3(s.adj[t], k);
        const int3 tv = s.tri[t], nv = s.tri[n];
        const int p = get3(tv, k);
        const int o = get3(nv, slot_other(nv, a, b));
        const int oa = stn_orient(s, p, o, a), ob = stn_orient(s, p, o, b);
        long long tail = (head + cnt) % qcap;
        if (oa != 0 && ob != 0 && oa != ob) {
            stn_flip(s, t, k);
            const int sp = stn_orient(s, x, y, p), so = stn_orient(s, x, y, o);
            if (sp != 0 && so != 0 && sp != so) {
                Q[2 * tail] = p;
                Q[2 * tail + 1] = o;
                cnt++;
            } else {
                if (nl >= lcap) {
                    at_or(s.err, MERR_SCRATCH);
                    return false;
                }
                L[2 * nl] = p;
                L[2 * nl + 1] = o;
                nl++;
            }
        } else {
            Q[2 * tail] = a;
            Q[2 * tail + 1] = b;
            cnt++;
        }
    }
    if (!stn_find_edge(s, x, y, t, k)) {
        at_or(s.err, MERR_CAVITY);
        return false;
    }
    stn_set_c(s, t, k);
    return stn_legalize(s, L, lcap, nl);
}

// records piece (a, b) as passing through Steiner vertex v
__host__ __device__ inline bool stn_add_sup(const Stn& s, int v, int a, int b)
{
    if (v < s.nv0) {  // a grid vertex inside a piece: the split pass missed it
        at_or(s.err, MERR_WALK);
        return false;
    }
    int2* sp = s.sup + (long long)STN_SUPS * (v - s.nv0);
    for (int j = 0; j < STN_SUPS; j++) {
        const int2 e = sp[j];
        if (e.x < 0) {
            sp[j] = make_int2(a, b);
            return true;
        }
        if ((e.x == a && e.y == b) || (e.x == b && e.y == a)) return true;
    }
    at_or(s.err, MERR_CAPACITY);
    return false;
}

// the piece a constrained edge (u, v) lies on, as two grid vertex ids
__host__ __device__ inline int2 stn_support(const Stn& s, int u, int v)
{
    if (u < s.nv0 && v < s.nv0) return make_int2(u, v);
    for (int pass = 0; pass < 2; pass++) {
        const int w = pass == 0 ? u : v, o = pass == 0 ? v : u;
        if (w < s.nv0) continue;
        const int2* sp = s.sup + (long long)STN_SUPS * (w - s.nv0);
        for (int j = 0; j < STN_SUPS; j++) {
            const int2 e = sp[j];
            if (e.x < 0) break;
            if (o < s.nv0 ? (o == e.x || o == e.y) : stn_orient(s, e.x, e.y, o) == 0) return e;
        }
    }
    return make_int2(-1, -1);
}

// Inserts deferred piece (a, b) (grid vertices).
__host__ __device__ inline bool stn_piece(const Stn& s, int a, int b)
{
    int cur = a;
    int* stk = s.scr;
    const long long cap_pairs = s.scr_cap / 2;
    for (int guard = 0; cur != b; guard++) {
        if (guard > (1 << 24)) {
            at_or(s.err, MERR_WALK);
            return false;
        }
        int t, w;
        const int r = stn_find_wedge(s, cur, b, t, w);
        if (r < 0) {
            at_or(s.err, MERR_WALK);
            return false;
        }
        const int3 tv = s.tri[t];
        if (r == 1) {  // the piece runs along edge (cur, u)
            const int u = get3(tv, nx3(w));
            if (u != b && !stn_add_sup(s, u, a, b)) return false;
            stn_set_c(s, t, pv3(w));
            cur = u;
            continue;
        }
        int rr = get3(tv, nx3(w)), ll = get3(tv, pv3(w));
        int tt = t, next = -1;
        for (int g = 0; g < (1 << 26); g++) {
            const int kk = slot_other(s.tri[tt], rr, ll);
            const int n = get3(s.adj[tt], kk);
            if (n < 0) break;
            if ((s.cflag[tt] >> kk) & 1) {
                // a constrained edge: split it where the piece crosses it
                const int2 e = stn_support(s, rr, ll);
                HPt X;
                if (e.x < 0 || !hpt_cross(s.pts[a], s.pts[b], s.pts[e.x], s.pts[e.y], X)) {
                    at_or(s.err, MERR_CAVITY);
                    return false;
                }
                const int k = s.ctr[STN_S];
                if (k >= s.scap) {
                    at_or(s.err, MERR_CAPACITY);
                    return false;
                }
                s.ctr[STN_S] = k + 1;
                const int sv = s.nv0 + k;
                s.hp[k] = X;
                int2* sp = s.sup + (long long)STN_SUPS * k;
                sp[0] = e;
                sp[1] = make_int2(a, b);
                for (int j = 2; j < STN_SUPS; j++) sp[j] = make_int2(-1, -1);
                long long top = 0;
                if (!stn_split(s, tt, kk, sv, stk, top) || !stn_legalize(s, stk, cap_pairs, top)) return false;
                next = sv;
                break;
            }
            const int3 nv = s.tri[n];
            const int qv = get3(nv, slot_other(nv, rr, ll));
            if (qv == b) {
                next = b;
                break;
            }
            const int o = stn_orient(s, a, b, qv);
            if (o == 0) {  // a crossing of two other pieces lies on this one
                if (!stn_add_sup(s, qv, a, b)) return false;
                next = qv;
                break;
            }
            if (o > 0) ll = qv;
            else rr = qv;
            tt = n;
        }
        if (next < 0) {
            at_or(s.err, MERR_WALK);
            return false;
        }
        if (!stn_recover(s, cur, next)) return false;
        cur = next;
    }
    return true;
}

}  // namespace gcdt
