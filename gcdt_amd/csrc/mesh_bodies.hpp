// Per-item bodies of the mesh kernels (point insertion, edge flips, mesh
// bookkeeping).  Each body handles one list item; the __global__ wrappers in
// mesh.hip add the wave-level list compaction around them.
//
// Mesh layout (identical to the reference's rawMesh layout, rawMesh.cu:5-22,
// 200-280):
//   pts[v]      int2 on the 2^30 grid; the 4 square corners are ids n..n+3
//   tri[t]      int3 vertex ids, counter-clockwise
//   adj[t]      int3, component k = neighbour across the edge opposite vertex k
//               (-1 = none)
//   sons[t]     insertion DAG: -1 leaf, >0 first of 3 children
//               (q,y,z),(q,z,x),(q,x,y) of an interior split by q, <0 minus the
//               first of 2 children (o,a,q),(o,q,b) of a split of edge (a,b)
//               of (o,a,b) by q
//   cflag[t]    bit k set = the edge opposite vertex k is a constraint
// Point state during insertion, loc[p]: -1 inserted / not pending, otherwise
// t*4+k with k==3 "strictly inside leaf t" and k<3 "on the edge opposite
// vertex k of leaf t".
#pragma once
#include "gcdt_common.hpp"
#include "predicates.hpp"

namespace gcdt {

// device atomics used by the bodies
__device__ inline unsigned long long at_max(unsigned long long* p, unsigned long long v) { return atomicMax(p, v); }
__device__ inline unsigned long long at_min(unsigned long long* p, unsigned long long v) { return atomicMin(p, v); }
__device__ inline int at_max(int* p, int v) { return atomicMax(p, v); }
__device__ inline int at_min(int* p, int v) { return atomicMin(p, v); }
__device__ inline int at_or(int* p, int v) { return atomicOr(p, v); }
__device__ inline int at_add(int* p, int v) { return atomicAdd(p, v); }
// Claim words are only ever changed by device-scope atomics (performed in L2);
// read them with an agent-scope atomic load so a stale L1 line is never used,
// also inside a persistent workgroup that loops over rounds.
__device__ inline unsigned long long ld_u64(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// COH: inside a persistent workgroup (atomic load); otherwise a kernel
// boundary separates the claims from this read and a plain load suffices.
template <bool COH>
__host__ __device__ inline unsigned long long ld_claim(const unsigned long long* p)
{
    if (COH) return ld_u64(p);
    return *p;
}

enum : int {
    MERR_CAPACITY = 1,     // triangle slots exhausted
    MERR_LOCATE = 2,       // a point was not found in any candidate triangle
    MERR_ADJ = 4,          // adjacency could not be resolved
    MERR_DUPLICATE = 8,    // a point coincides with a vertex
    MERR_CROSSING = 16,    // two constraints cross
    MERR_WALK = 32,        // a constraint walk left the mesh
    MERR_CAVITY = 64,      // cavity retriangulation inconsistency
    MERR_SCRATCH = 128,    // cavity scratch exhausted
    MERR_INVALID = 256,    // input outside the contract (coordinates, ids)
    MERR_INTERNAL_COUNT = 512,  // per-instance triangle count mismatch
};

struct Mesh {
    const int2* pts;
    int3* tri;
    int3* adj;
    int* sons;
    int* cflag;
};

__host__ __device__ inline int& A3(const Mesh& m, int t, int k) { return ((int*)m.adj)[3 * (long long)t + k]; }

// ------------------------------------------------------------------ insertion

struct Ins {
    Mesh m;
    int* loc;
    unsigned long long* claim;
    unsigned long long* pkey;  // per point: its claim key of the current round
    int cap;
    int pivot;                 // 0: random among the triangle's points, 1: central ones first
    int* err;
};

// Claim priority: newer epoch first, then on-edge points (they need two
// triangles), then the point nearest the triangle's centroid (balanced
// splits: O(log m) rounds for m points in a triangle, and well-shaped children
// that need few flips later), then the id.  Layout: epoch 20 bits | on-edge 1 |
// closeness 11 | id 32.
__host__ __device__ inline unsigned closeness_rank(int2 A, int2 B, int2 C, int2 P, int p)
{
    long long dx = 3LL * P.x - ((long long)A.x + B.x + C.x);
    long long dy = 3LL * P.y - ((long long)A.y + B.y + C.y);
    long long d = (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy);
    long long size = (long long)(A.x > B.x ? A.x - B.x : B.x - A.x) + (A.y > B.y ? A.y - B.y : B.y - A.y) +
                     (long long)(B.x > C.x ? B.x - C.x : C.x - B.x) + (B.y > C.y ? B.y - C.y : C.y - B.y) +
                     (long long)(C.x > A.x ? C.x - A.x : A.x - C.x) + (C.y > A.y ? C.y - A.y : A.y - C.y);
    // class 3: well inside the triangle ... 0: near its boundary; random within a class
    unsigned cls = d * 8 < size ? 3u : (d * 4 < size ? 2u : (d * 2 < size ? 1u : 0u));
    return (cls << 9) | (hash32((uint32_t)p) & 511u);
}

__host__ __device__ inline unsigned long long ins_key(int p, int onedge, unsigned close, unsigned epoch)
{
    return ((unsigned long long)(epoch & 0xfffffu) << 44) | ((unsigned long long)onedge << 43) |
           ((unsigned long long)(close & 2047u) << 32) | (unsigned)p;
}

// A point on the edge opposite slot k < 3 is ranked against that edge alone:
// its two triangles may record it in either, and both must give one key.
__host__ __device__ inline unsigned closeness_rank_edge(int2 U, int2 V, int2 P, int p)
{
    long long dx = 2LL * P.x - ((long long)U.x + V.x);
    long long dy = 2LL * P.y - ((long long)U.y + V.y);
    long long d = (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy);
    long long size = 2LL * ((U.x > V.x ? (long long)U.x - V.x : (long long)V.x - U.x) +
                            (U.y > V.y ? (long long)U.y - V.y : (long long)V.y - U.y));
    unsigned cls = d * 8 < size ? 3u : (d * 4 < size ? 2u : (d * 2 < size ? 1u : 0u));
    return (cls << 9) | (hash32((uint32_t)p) & 511u);
}

__host__ __device__ inline unsigned long long ins_key_of(const Ins& s, int p, int t, int k, unsigned epoch)
{
    unsigned close;
    if (s.pivot == 1) {
        int3 tv = s.m.tri[t];
        close = k < 3 ? closeness_rank_edge(s.m.pts[get3(tv, nx3(k))], s.m.pts[get3(tv, pv3(k))], s.m.pts[p], p)
                      : closeness_rank(s.m.pts[tv.x], s.m.pts[tv.y], s.m.pts[tv.z], s.m.pts[p], p);
    } else {
        close = hash32((uint32_t)p) & 2047u;
    }
    return ins_key(p, k < 3, close, epoch);
}

// Computes p's key for this round (cached in pkey) and returns its triangle
// (the caller performs the claim so a wave can combine claims on one triangle);
// -1 when p is not pending.  *t2 receives the second triangle of an on-edge
// point, else -1.
__host__ __device__ inline int ins_key_round(const Ins& s, int p, unsigned epoch, unsigned long long* key, int* t2)
{
    int l = s.loc[p];
    *t2 = -1;
    if (l < 0) return -1;
    int t = l >> 2, k = l & 3;
    *key = ins_key_of(s, p, t, k, epoch);
    s.pkey[p] = *key;
    if (k < 3) *t2 = A3(s.m, t, k);
    return t;
}

__host__ __device__ inline void ins_claim(const Ins& s, int p, unsigned epoch)
{
    unsigned long long key;
    int t2;
    int t = ins_key_round(s, p, epoch, &key, &t2);
    if (t < 0) return;
    at_max(&s.claim[t], key);
    if (t2 >= 0) at_max(&s.claim[t2], key);
}

// slots needed by p if it won every triangle it claimed (3 interior, 4 edge)
template <bool COH = true>
__host__ __device__ inline int ins_need(const Ins& s, int p, unsigned epoch)
{
    int l = s.loc[p];
    if (l < 0) return 0;
    int t = l >> 2, k = l & 3;
    unsigned long long key = s.pkey[p];
    if (ld_claim<COH>(&s.claim[t]) != key) return 0;
    if (k == 3) return 3;
    return ld_claim<COH>(&s.claim[A3(s.m, t, k)]) == key ? 4 : 0;
}

// Split the triangle(s) of winner p into new slots [base, base+need).
// Reference: addNewTriangle_step2, rawMesh.cu:200-280 (same child layout).
__host__ __device__ inline void ins_split(const Ins& s, int p, int base)
{
    const Mesh& m = s.m;
    int l = s.loc[p];
    int t = l >> 2, k = l & 3;
    int3 tv = m.tri[t];
    int3 ta = m.adj[t];
    if (k == 3) {
        int a = tv.x, b = tv.y, c = tv.z;
        m.tri[base] = make_int3(p, b, c);
        m.adj[base] = make_int3(adj_mark(ta.x), base + 1, base + 2);
        m.tri[base + 1] = make_int3(p, c, a);
        m.adj[base + 1] = make_int3(adj_mark(ta.y), base + 2, base);
        m.tri[base + 2] = make_int3(p, a, b);
        m.adj[base + 2] = make_int3(adj_mark(ta.z), base, base + 1);
        m.sons[t] = base;
    } else {
        int n = get3(ta, k);
        int o = get3(tv, k), a = get3(tv, nx3(k)), b = get3(tv, pv3(k));
        int at_a = get3(ta, nx3(k));  // across (b,o)
        int at_b = get3(ta, pv3(k));  // across (o,a)
        int3 nv = m.tri[n];
        int3 na = m.adj[n];
        int j = slot_other(nv, a, b);
        int o2 = get3(nv, j);         // n = (o2, b, a) rotated
        int an_b = get3(na, nx3(j));  // across (a,o2)
        int an_a = get3(na, pv3(j));  // across (o2,b)
        // t side: (o,a,p) and (o,p,b)
        m.tri[base] = make_int3(o, a, p);
        m.adj[base] = make_int3(base + 3, base + 1, adj_mark(at_b));
        m.tri[base + 1] = make_int3(o, p, b);
        m.adj[base + 1] = make_int3(base + 2, adj_mark(at_a), base);
        // n side: (o2,b,p) and (o2,p,a)
        m.tri[base + 2] = make_int3(o2, b, p);
        m.adj[base + 2] = make_int3(base + 1, base + 3, adj_mark(an_a));
        m.tri[base + 3] = make_int3(o2, p, a);
        m.adj[base + 3] = make_int3(base, adj_mark(an_b), base + 2);
        m.sons[t] = -base;
        m.sons[n] = -(base + 2);
    }
    s.loc[p] = -1;
}

// Resolve the outer adjacency of new slot c after all splits of the round:
// a marked entry names the parent's old neighbour m; if m was split in the
// same round, the right child of m is the one holding both edge vertices,
// otherwise m keeps its slot and gets a back-pointer to c.
// (The reference instead repairs leaves once at the end by descending the
// DAG, updateAdjTris_new, rawMesh.cu:590-676.)
__host__ __device__ inline void ins_fix(const Ins& s, int c)
{
    const Mesh& m = s.m;
    int3 tv = m.tri[c];
    int3 ta = m.adj[c];
    for (int k = 0; k < 3; k++) {
        int a = get3(ta, k);
        if (!adj_is_mark(a)) continue;
        int nb = adj_unmark(a);
        int u = get3(tv, nx3(k)), v = get3(tv, pv3(k));
        int sn = m.sons[nb];
        if (sn == -1) {
            set3(ta, k, nb);
            A3(m, nb, slot_other(m.tri[nb], u, v)) = c;
        } else {
            int c0 = sn > 0 ? sn : -sn, nc = sn > 0 ? 3 : 2;
            int q = -1;
            for (int x = c0; x < c0 + nc; x++)
                if (has2(m.tri[x], u, v)) { q = x; break; }
            if (q < 0) at_or(s.err, MERR_ADJ);
            set3(ta, k, q);
        }
    }
    m.adj[c] = ta;
}

// Closed point-in-triangle: returns -1 outside, 3 strictly inside, or the slot
// k of the edge (opposite vertex k) that p lies on.
__host__ __device__ inline int classify(int2 A, int2 B, int2 C, int2 P)
{
    int o0 = orient2d(B, C, P);
    if (o0 < 0) return -1;
    int o1 = orient2d(C, A, P);
    if (o1 < 0) return -1;
    int o2 = orient2d(A, B, P);
    if (o2 < 0) return -1;
    int z = (o0 == 0) + (o1 == 0) + (o2 == 0);
    if (z == 0) return 3;
    if (z > 1) return -2;  // p is a vertex
    return o0 == 0 ? 0 : (o1 == 0 ? 1 : 2);
}

// After the splits: move a still-pending point into the child of its split
// triangle that contains it.  Returns 1 if p stays pending.
// Reference: scatterPoints/scatterPoint, rawMesh.cu:338-463.
__host__ __device__ inline int ins_relocate(const Ins& s, int p)
{
    const Mesh& m = s.m;
    int l = s.loc[p];
    if (l < 0) return 0;
    int t = l >> 2;
    int sn = m.sons[t];
    if (sn == -1) return 1;  // untouched (an on-edge point keeps its edge: see ins_fix)
    int c0 = sn > 0 ? sn : -sn, nc = sn > 0 ? 3 : 2;
    int2 P = m.pts[p];
    for (int c = c0; c < c0 + nc; c++) {
        int3 tv = m.tri[c];
        int r = classify(m.pts[tv.x], m.pts[tv.y], m.pts[tv.z], P);
        if (r >= 0) {
            s.loc[p] = c * 4 + r;
            return 1;
        }
        if (r == -2) {
            at_or(s.err, MERR_DUPLICATE);
            s.loc[p] = -1;
            return 0;
        }
    }
    at_or(s.err, MERR_LOCATE);
    s.loc[p] = -1;
    return 0;
}

// ------------------------------------------------------------------ flips

struct Flip {
    Mesh m;
    unsigned long long* claim;
    int* stamp;    // epoch of the last flip of this triangle
    int* partner;  // the other triangle of that flip
    int* cand;     // proposed edge slot, -1 none
    int* mark;     // epoch in which the triangle is in the active list
    int* err;
};

__host__ __device__ inline int opposite_vertex(const Mesh& m, int n, int u, int v)
{
    int3 nv = m.tri[n];
    return get3(nv, slot_other(nv, u, v));
}

// Test the edges of t; propose the first edge that is not locally Delaunay
// (strict inCircle > 0, never a constrained or boundary edge) by claiming
// both of its triangles with the edge key.
// Reference: initNeedFlipFE_new, kernel.cu:2198-2228.
// An edge whose two triangles are both active is tested by the lower id only
// (all_active: every triangle is active this epoch).  Every illegal edge keeps
// an active triangle: the lower one proposes an edge, so it is re-queued or
// flipped.
__host__ __device__ inline bool flip_test(const Flip& f, int t, unsigned epoch, bool all_active, bool single = true)
{
    const Mesh& m = f.m;
    int3 tv = m.tri[t];
    int3 ta = m.adj[t];
    int cf = m.cflag ? m.cflag[t] : 0;
    int2 A = m.pts[tv.x], B = m.pts[tv.y], C = m.pts[tv.z];
    int cand = -1;
    for (int k = 0; k < 3; k++) {
        if ((cf >> k) & 1) continue;
        int n = get3(ta, k);
        if (n < 0) continue;
        if (single && n < t && (all_active || f.mark[n] == (int)epoch)) continue;
        int o = opposite_vertex(m, n, get3(tv, nx3(k)), get3(tv, pv3(k)));
        if (incircle(A, B, C, m.pts[o]) > 0) {
            cand = k;
            unsigned long long key = ((unsigned long long)epoch << 32) | (unsigned)(3 * t + k);
            at_max(&f.claim[t], key);
            at_max(&f.claim[n], key);
            break;
        }
    }
    f.cand[t] = cand;
    return cand >= 0;
}

// Flip t's proposed edge if it won both triangles.  Returns 0 no proposal,
// 1 lost (t must be re-tested), 2 flipped (t and *nout changed), 3 lost to a
// proposal that wins and flips t (its winner re-queues t, so t does not).
// t=(p,u,v), n=(o,v,u)  ->  t=(p,u,o), n=(p,o,v); outer edges are marked
// for flip_fix.  Reference: flipEdgeFE / multiFlipFE, kernel.cu:2681-2710,3548.
template <bool COH = true>
__host__ __device__ inline int flip_apply(const Flip& f, int t, unsigned epoch, int* nout)
{
    const Mesh& m = f.m;
    int k = f.cand[t];
    if (k < 0) return 0;
    int3 tv = m.tri[t];
    int3 ta = m.adj[t];
    int n = get3(ta, k);
    unsigned long long key = ((unsigned long long)epoch << 32) | (unsigned)(3 * t + k);
    unsigned long long ct = ld_claim<COH>(&f.claim[t]);
    if (ct != key || ld_claim<COH>(&f.claim[n]) != key) {
        if (ct != key) {
            // ct was proposed by triangle t2 for the edge it shares with t; t2
            // wins (and flips t) iff it also holds its own claim
            int t2 = (int)((ct & 0xffffffffu) / 3);
            if ((ct >> 32) == epoch && ld_claim<COH>(&f.claim[t2]) == ct) return 3;
        }
        return 1;
    }
    int p = get3(tv, k), u = get3(tv, nx3(k)), v = get3(tv, pv3(k));
    int3 nv = m.tri[n];
    int3 na = m.adj[n];
    int j = slot_other(nv, u, v);
    int o = get3(nv, j);
    int nb_uo = get3(na, nx3(j)), nb_ov = get3(na, pv3(j));
    int nb_pu = get3(ta, pv3(k)), nb_vp = get3(ta, nx3(k));
    int cft = m.cflag ? m.cflag[t] : 0, cfn = m.cflag ? m.cflag[n] : 0;
    int c_uo = (cfn >> nx3(j)) & 1, c_ov = (cfn >> pv3(j)) & 1;
    int c_pu = (cft >> pv3(k)) & 1, c_vp = (cft >> nx3(k)) & 1;
    m.tri[t] = make_int3(p, u, o);
    m.adj[t] = make_int3(adj_mark(nb_uo), n, adj_mark(nb_pu));
    m.tri[n] = make_int3(p, o, v);
    m.adj[n] = make_int3(adj_mark(nb_ov), adj_mark(nb_vp), t);
    if (m.cflag) {
        m.cflag[t] = c_uo | (c_pu << 2);
        m.cflag[n] = c_ov | (c_vp << 1);
    }
    f.stamp[t] = (int)epoch;
    f.stamp[n] = (int)epoch;
    f.partner[t] = n;
    f.partner[n] = t;
    *nout = n;
    return 2;
}

// Resolve the marked outer adjacency of a triangle flipped this epoch.
__host__ __device__ inline void flip_fix(const Flip& f, int t, unsigned epoch)
{
    const Mesh& m = f.m;
    int3 tv = m.tri[t];
    int3 ta = m.adj[t];
    for (int k = 0; k < 3; k++) {
        int a = get3(ta, k);
        if (!adj_is_mark(a)) continue;
        int nb = adj_unmark(a);
        int u = get3(tv, nx3(k)), v = get3(tv, pv3(k));
        if (f.stamp[nb] == (int)epoch) {
            int q = has2(m.tri[nb], u, v) ? nb : f.partner[nb];
            set3(ta, k, q);
        } else {
            set3(ta, k, nb);
            A3(m, nb, slot_other(m.tri[nb], u, v)) = t;
        }
    }
    m.adj[t] = ta;
}

// ---------------------------------------------------------------- packed flip records

// One 64-byte record per triangle for flip rounds over a whole mesh.  A test
// or a flip reads a neighbour's vertices, neighbours, vertex coordinates,
// flip bookkeeping and claim word from one record -- one line -- where the
// array-per-field layout gathers from up to eight places.
struct FRec {
    int3 v;        // vertices, counter-clockwise
    int3 a;        // neighbour across the edge opposite v[k]; marked while a flip's outer edges settle
    int2 p[3];     // coordinates of v[0..2]
    int meta;      // bits 0-2 constrained edges (bit k = edge opposite v[k]); 3-4 proposed edge (3 none);
                   // 5-6 slot of the last flip's partner (3 none); 7-31 epoch in which it is active
    int stamp;     // epoch of its last flip
    unsigned long long claim;  // highest proposal key of the round (atomicMax)
};
static_assert(sizeof(FRec) == 64, "flip record is one 64-byte line (arrays of records start 256-byte aligned)");

constexpr int FREC_NOCAND = 3 << 3;
constexpr int FREC_NOPARTNER = 3 << 5;
constexpr unsigned FREC_MARK_MASK = 0x1FFFFFFu;  // 25-bit epochs (a solve runs far fewer rounds)

__host__ __device__ inline int rec_cand(int meta) { return (meta >> 3) & 3; }
__host__ __device__ inline int rec_pslot(int meta) { return (meta >> 5) & 3; }
__host__ __device__ inline bool rec_marked(int meta, unsigned epoch)
{
    return ((unsigned)meta >> 7) == (epoch & FREC_MARK_MASK);
}
__host__ __device__ inline int rec_with_mark(int meta, int epoch)
{
    return (meta & 0x7F) | (int)(((unsigned)epoch & FREC_MARK_MASK) << 7);
}

// Whole-record load and store as four 16-byte accesses (a struct copy of an
// over-aligned local is otherwise staged through the stack).
__host__ __device__ inline FRec rec_load(const FRec* p)
{
    const int4* q = reinterpret_cast<const int4*>(p);
    int4 a = q[0];
    GCDT_MEM_POINT();
    int4 b = q[1];
    GCDT_MEM_POINT();
    int4 c = q[2];
    GCDT_MEM_POINT();
    int4 d = q[3];
    FRec r;
    r.v = make_int3(a.x, a.y, a.z);
    r.a = make_int3(a.w, b.x, b.y);
    r.p[0] = make_int2(b.z, b.w);
    r.p[1] = make_int2(c.x, c.y);
    r.p[2] = make_int2(c.z, c.w);
    r.meta = d.x;
    r.stamp = d.y;
    r.claim = (unsigned long long)(unsigned)d.z | ((unsigned long long)(unsigned)d.w << 32);
    return r;
}

__host__ __device__ inline void rec_store(FRec* p, const int3& v, const int3& a, const int2& p0, const int2& p1,
                                          const int2& p2, int meta, int stamp, unsigned long long claim)
{
    int4* q = reinterpret_cast<int4*>(p);
    q[0] = make_int4(v.x, v.y, v.z, a.x);
    GCDT_MEM_POINT();
    q[1] = make_int4(a.y, a.z, p0.x, p0.y);
    GCDT_MEM_POINT();
    q[2] = make_int4(p1.x, p1.y, p2.x, p2.y);
    GCDT_MEM_POINT();
    q[3] = make_int4(meta, stamp, (int)(unsigned)claim, (int)(unsigned)(claim >> 32));
}

// coordinate k of a record by selects over values (a runtime index into the
// local copy, or selects through a reference, would put it on the stack)
__host__ __device__ inline int2 sel_pt(int2 p0, int2 p1, int2 p2, int k)
{
    return make_int2(k == 0 ? p0.x : (k == 1 ? p1.x : p2.x), k == 0 ? p0.y : (k == 1 ? p1.y : p2.y));
}
#define rec_pt(r, k) sel_pt((r).p[0], (r).p[1], (r).p[2], (k))

struct FlipP {
    FRec* r;
    int* err;
};

// flip_test over records (same rule as flip_test); true if t proposed.
__host__ __device__ inline bool flip_test_p(const FlipP& f, int t, unsigned epoch, bool all_active, bool single = true)
{
    const FRec r = rec_load(&f.r[t]);
    const int cf = r.meta & 7;
    const int ps = rec_pslot(r.meta);
    const int partner = (r.stamp == (int)epoch - 1 && ps < 3) ? get3(r.a, ps) : -1;
    int cand = 3;
    for (int k = 0; k < 3; k++) {
        if ((cf >> k) & 1) continue;
        int n = get3(r.a, k);
        if (n < 0) continue;
        // the diagonal of last round's flip is locally Delaunay (it replaced a
        // strictly illegal edge, and neither triangle has changed since)
        if (n == partner) continue;
        bool lower = single && n < t;
        if (lower && all_active) continue;
        const FRec rn = rec_load(&f.r[n]);
        if (lower && rec_marked(rn.meta, epoch)) continue;
        int j = slot_other(rn.v, get3(r.v, nx3(k)), get3(r.v, pv3(k)));
        if (incircle(r.p[0], r.p[1], r.p[2], rec_pt(rn, j)) > 0) {
            cand = k;
            unsigned long long key = ((unsigned long long)epoch << 32) | (unsigned)(3 * t + k);
            at_max(&f.r[t].claim, key);
            at_max(&f.r[n].claim, key);
            break;
        }
    }
    const int meta = (r.meta & ~(3 << 3)) | (cand << 3);
    if (meta != r.meta) f.r[t].meta = meta;
    return cand != 3;
}

// Where a packed flip round keeps its claims: the claim word of each record
// (RecClaims), or a table in LDS for the short rounds of one persistent
// workgroup (mesh.hip).  get<COH> returns t's claim; `loaded` is the word
// already read with t's record, used when no re-read is needed.
struct RecClaims {
    __host__ __device__ void max(const FlipP& f, int t, unsigned long long key) const { at_max(&f.r[t].claim, key); }
    template <bool COH>
    __host__ __device__ unsigned long long get(const FlipP& f, int t, unsigned long long loaded) const
    {
        return COH ? ld_claim<COH>(&f.r[t].claim) : loaded;
    }
    // t's claim when its record has not been read
    template <bool COH>
    __host__ __device__ unsigned long long load(const FlipP& f, int t) const
    {
        return ld_claim<COH>(&f.r[t].claim);
    }
};

// flip_fix of the previous round and flip_test of this one in a single pass
// over t's record and its neighbours' (each neighbour record is loaded once
// for both).  Runs over the round's list, which holds every triangle flipped
// in round epoch - 1:
// * if t flipped last round (stamp = epoch - 1), its two slots other than the
//   partner's name the old neighbours nb of the flipped pair: if nb flipped
//   too, the edge now belongs to nb or to nb's partner; otherwise t writes the
//   back-pointer into nb.  Only slots that change are stored (most keep their
//   value), so back-pointers written by other triangles into t are never lost
//   and unchanged lines stay clean.
// * any other entry may still name a triangle that flipped last round (its
//   fix runs in this same pass): a neighbour record without both edge
//   vertices is resolved to its partner of that flip.
// By the end of the launch every entry and back-pointer is settled, as after
// a separate fix launch; the last round's flips are settled by flip_fix_p.
template <class CL = RecClaims>
__host__ __device__ inline bool flip_fixtest_p(const FlipP& f, int t, unsigned epoch, bool all_active,
                                               bool single = true, const CL& cl = CL())
{
    const FRec r = rec_load(&f.r[t]);
    const int prev = (int)epoch - 1;
    const int cf = r.meta & 7;
    const int ps = rec_pslot(r.meta);
    const bool flipped = r.stamp == prev && ps < 3;
    const int partner = flipped ? get3(r.a, ps) : -1;
    // every neighbour record the pass can need, requested together before any
    // store (three line reads in flight instead of a chain of three)
    FRec nr[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int av = get3(r.a, k);
        const bool pend = flipped && k != ps && av >= 0;
        const bool tst = !((cf >> k) & 1) && av >= 0 && av != partner && !(single && av < t && all_active);
        if (pend || tst) nr[k] = rec_load(&f.r[av]);
    }
    int cand = 3;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int av = get3(r.a, k);
        const int u = get3(r.v, nx3(k)), v = get3(r.v, pv3(k));
        int n = av;
        FRec rn = nr[k];
        bool have = false;
        if (flipped && k != ps && av >= 0) {
            // an outer slot of last round's flip: it names the old neighbour
            if (rn.stamp == prev) {
                if (!has2(rn.v, u, v)) {
                    n = get3(rn.a, rec_pslot(rn.meta));
                    rn = rec_load(&f.r[n]);
                    (&f.r[t].a.x)[k] = n;
                }
            } else {
                // the outer edge that kept its triangle already points here
                const int j = slot_other(rn.v, u, v);
                if (get3(rn.a, j) != t) (&f.r[av].a.x)[j] = t;
            }
            have = true;
        }
        if (cand != 3 || ((cf >> k) & 1) || n < 0 || n == partner) continue;
        if (single && n < t && all_active) continue;
        if (!have) {
            if (!all_active && !has2(rn.v, u, v)) {
                // n flipped last round and the edge went to its partner
                if (rn.stamp != prev) {
                    at_or(f.err, MERR_ADJ);
                    continue;
                }
                n = get3(rn.a, rec_pslot(rn.meta));
                rn = rec_load(&f.r[n]);
            }
        }
        if (single && n < t && rec_marked(rn.meta, epoch)) continue;
        const int j = slot_other(rn.v, u, v);
        if (incircle(r.p[0], r.p[1], r.p[2], rec_pt(rn, j)) > 0) {
            cand = k;
            const unsigned long long key = ((unsigned long long)epoch << 32) | (unsigned)(3 * t + k);
            cl.max(f, t, key);
            cl.max(f, n, key);
        }
    }
    // most tested triangles propose nothing and already say so: a store would
    // only dirty the line
    const int meta = (r.meta & ~(3 << 3)) | (cand << 3);
    if (meta != r.meta) f.r[t].meta = meta;
    return cand != 3;
}

// flip_apply over records: same claims protocol and return codes.  A flip
// rewrites both records whole -- keeping the claim word, which losers may be
// reading, at the winning key -- with the new vertices, coordinates,
// constraint bits, partner slot, stamp = epoch and mark = epoch + 1 (both are
// queued).  The two outer slots of each keep the old neighbours' ids; the
// stamp marks them for settling by the next round's test (flip_fixtest_p) or
// by flip_fix_p.
//
// When both triangles of an edge test it (the persistent tail, or
// single_tester 0), a proposal can lose its own triangle to another one that
// flips it in this same pass.  Its record may then be read while that flip
// stores it, 16 bytes at a time: some parts new, some old.  Only the claim
// word is the same in every such view (the winning key), so nothing else of
// the record -- above all no neighbour id, which may be -1 in the new parts
// -- is used before the claim says t is still its own.
template <bool COH = true, class CL = RecClaims>
__host__ __device__ inline int flip_apply_p(const FlipP& f, int t, unsigned epoch, int* nout, const CL& cl = CL())
{
    const FRec r = rec_load(&f.r[t]);
    const int k = rec_cand(r.meta);
    if (k == 3) return 0;
    const unsigned long long key = ((unsigned long long)epoch << 32) | (unsigned)(3 * t + k);
    const unsigned long long ct = cl.template get<COH>(f, t, r.claim);
    if (ct != key) {
        int t2 = (int)((ct & 0xffffffffu) / 3);
        if ((ct >> 32) == epoch && cl.template load<COH>(f, t2) == ct) return 3;
        return 1;
    }
    // t holds its own claim, so no other flip writes its record in this pass
    const int n = get3(r.a, k);
    const FRec rn = rec_load(&f.r[n]);
    const unsigned long long cn = cl.template get<COH>(f, n, rn.claim);
    if (cn != key) return 1;
    const int kn = nx3(k), kp = pv3(k);
    const int p = get3(r.v, k), u = get3(r.v, kn), v = get3(r.v, kp);
    const int j = slot_other(rn.v, u, v);
    const int jn = nx3(j), jp = pv3(j);
    const int o = get3(rn.v, j);
    const int cft = r.meta & 7, cfn = rn.meta & 7;
    const int2 pp = rec_pt(r, k), pu = rec_pt(r, kn), pv = rec_pt(r, kp), po = rec_pt(rn, j);
    const int mk = (int)epoch + 1;
    // t = (p,u,v), n = (o,v,u)  ->  t = (p,u,o) [partner in slot 1], n = (p,o,v) [partner in slot 2]
    rec_store(&f.r[t], make_int3(p, u, o), make_int3(get3(rn.a, jn), n, get3(r.a, kp)), pp, pu, po,
              rec_with_mark(((cfn >> jn) & 1) | (((cft >> kp) & 1) << 2) | FREC_NOCAND | (1 << 5), mk), (int)epoch,
              key);
    rec_store(&f.r[n], make_int3(p, o, v), make_int3(get3(rn.a, jp), get3(r.a, kn), t), pp, po, pv,
              rec_with_mark(((cfn >> jp) & 1) | (((cft >> kn) & 1) << 1) | FREC_NOCAND | (2 << 5), mk), (int)epoch,
              key);
    *nout = n;
    return 2;
}

// flip_fix over records: settles the outer slots of t, flipped in round
// epoch, and the back-pointers of its unflipped neighbours.
__host__ __device__ inline void flip_fix_p(const FlipP& f, int t, unsigned epoch)
{
    const FRec r = rec_load(&f.r[t]);
    const int ps = rec_pslot(r.meta);
    if (r.stamp != (int)epoch || ps == 3) return;
    for (int k = 0; k < 3; k++) {
        const int nb = get3(r.a, k);
        if (k == ps || nb < 0) continue;
        const int u = get3(r.v, nx3(k)), v = get3(r.v, pv3(k));
        const FRec rb = rec_load(&f.r[nb]);
        if (rb.stamp == (int)epoch) {
            if (!has2(rb.v, u, v)) (&f.r[t].a.x)[k] = get3(rb.a, rec_pslot(rb.meta));
        } else {
            const int j = slot_other(rb.v, u, v);
            if (get3(rb.a, j) != t) (&f.r[nb].a.x)[j] = t;
        }
    }
}

}  // namespace gcdt
