/*
 * cdt_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Serial, exact-predicate CPU restatement of the constrained Delaunay
 * triangulation that the reference computes on the GPU.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker or as the timed CPU baseline; the product
 * path (gcdt_amd) never calls it.
 *
 * What it restates (reference file:line):
 *   - orient2d sign convention      include/math_base.cuh:68-76
 *     (+1 counter-clockwise, 0 collinear, -1 clockwise; int64 products)
 *   - inCircle 3x3 lifted determinant, six-term expansion with 128-bit sums
 *                                   include/math_base.cuh:267-328
 *     (+1 means p strictly inside the circumcircle of CCW a,b,c)
 *   - the bounding square: corners (0,0),(R,0),(R,R),(0,R), R = 2^30-1, at
 *     vertex ids N..N+3, first split along the anti-diagonal into the
 *     triangles (N,N+1,N+3) and (N+1,N+2,N+3)       rawMesh.cu:5-22
 *   - the flip rule: an edge is flipped iff inCircle(face, opposite) > 0
 *     (strict), square-boundary and constrained edges never flip
 *                                   kernel.cu:2198-2228
 *   - constraints are vertex-id pairs; an edge that already exists is kept
 *     and protected                  addConstraints.cu:809-811, kernel.cu:1639-1662
 *   Deliberate deviation (same as the GPU product): a constraint that passes
 *   exactly through other vertices is split at them instead of being dropped
 *   (the reference drops it, addConstraints.cu:962-965).
 *
 * Algorithm: incremental Lawson insertion in Hilbert order with a visibility
 * walk, then per-segment cavity retriangulation (each side of the segment is
 * a pseudo-polygon triangulated by the recursive Delaunay-vertex rule), then
 * constrained Lawson flips as a safety net.  In general position the CDT is
 * unique, so this oracle and any correct GPU algorithm produce the same
 * triangle set.  The exact validator below checks any triangulation against
 * the definition (tiling, constraint chains present, every unconstrained edge
 * locally Delaunay), which pins non-unique (degenerate) cases.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef long long i64;

#define GRID_R ((1 << 30) - 1)

enum {
    OR_OK = 0,
    OR_ERR_ALLOC = 1,
    OR_ERR_DUPLICATE = 2,
    OR_ERR_RANGE = 3,
    OR_ERR_CONSTRAINT = 4,
    OR_ERR_CROSSING = 5,
    OR_ERR_INTERNAL = 6
};

/* ---------------------------------------------------------------- predicates */

/* sign((a-c) x (b-c)), math_base.cuh:68-76 */
int oracle_orient2d(i64 ax, i64 ay, i64 bx, i64 by, i64 cx, i64 cy)
{
    i64 o1 = (ax - cx) * (by - cy);
    i64 o2 = (ay - cy) * (bx - cx);
    return o1 > o2 ? 1 : (o1 < o2 ? -1 : 0);
}

/* six-term expansion of | ap.x ap.y |ap|^2 ; bp.. ; cp.. |, math_base.cuh:267-328 */
int oracle_incircle(i64 ax, i64 ay, i64 bx, i64 by, i64 cx, i64 cy, i64 px, i64 py)
{
    i64 apx = ax - px, apy = ay - py;
    i64 bpx = bx - px, bpy = by - py;
    i64 cpx = cx - px, cpy = cy - py;
    i128 aa = (i128)apx * apx + (i128)apy * apy;
    i128 bb = (i128)bpx * bpx + (i128)bpy * bpy;
    i128 cc = (i128)cpx * cpx + (i128)cpy * cpy;
    i128 sum = 0;
    sum += (i128)(apx * bpy) * cc;
    sum -= (i128)(apx * cpy) * bb;
    sum -= (i128)(apy * bpx) * cc;
    sum += (i128)(apy * cpx) * bb;
    sum += (i128)(bpx * cpy) * aa;
    sum -= (i128)(cpx * bpy) * aa;
    return sum > 0 ? 1 : (sum < 0 ? -1 : 0);
}

/* ---------------------------------------------------------------- mesh */

typedef struct {
    const int32_t* P; /* 2 * nv coordinates */
    int nv;
    int* V;           /* 3 per triangle, CCW */
    int* A;           /* neighbour across the edge opposite slot k, -1 = none */
    unsigned char* C; /* 1 if the edge opposite slot k is constrained */
    int T, cap;
    int* vt;          /* vertex -> some incident triangle */
    int* stk;         /* legalisation stack of (tri*3+slot) */
    int stk_n, stk_cap;
    i64 flips;
    int general;      /* 1: after a flip re-test all four outer edges */
    i64 cavity_flips; /* flips needed after cavity retriangulation (expected 0) */
} mesh_t;

static inline int ORI(const mesh_t* m, int a, int b, int c)
{
    const int32_t* P = m->P;
    return oracle_orient2d(P[2 * a], P[2 * a + 1], P[2 * b], P[2 * b + 1], P[2 * c], P[2 * c + 1]);
}

static inline int INC(const mesh_t* m, int a, int b, int c, int d)
{
    const int32_t* P = m->P;
    return oracle_incircle(P[2 * a], P[2 * a + 1], P[2 * b], P[2 * b + 1], P[2 * c], P[2 * c + 1],
                           P[2 * d], P[2 * d + 1]);
}

#define NX(k) (((k) + 1) % 3)
#define PV(k) (((k) + 2) % 3)

static int slot_of_nb(const mesh_t* m, int t, int nb)
{
    for (int k = 0; k < 3; k++)
        if (m->A[3 * t + k] == nb) return k;
    return -1;
}

static void set_tri(mesh_t* m, int t, int a, int b, int c)
{
    m->V[3 * t] = a;
    m->V[3 * t + 1] = b;
    m->V[3 * t + 2] = c;
    m->vt[a] = t;
    m->vt[b] = t;
    m->vt[c] = t;
}

static int push(mesh_t* m, int t, int k)
{
    if (m->stk_n == m->stk_cap) {
        int nc = m->stk_cap * 2 + 64;
        int* s = (int*)realloc(m->stk, sizeof(int) * (size_t)nc);
        if (!s) return OR_ERR_ALLOC;
        m->stk = s;
        m->stk_cap = nc;
    }
    m->stk[m->stk_n++] = 3 * t + k;
    return OR_OK;
}

/* Flip the edge opposite slot k of t.  With t=(p,u,v) and neighbour n=(o,v,u)
 * the quad p,u,o,v becomes t=(p,u,o), n=(p,o,v). */
static void flip(mesh_t* m, int t, int k)
{
    int* V = m->V;
    int* A = m->A;
    unsigned char* C = m->C;
    int p = V[3 * t + k], u = V[3 * t + NX(k)], v = V[3 * t + PV(k)];
    int n = A[3 * t + k];
    int j = slot_of_nb(m, n, t);
    int o = V[3 * n + j];
    int nb_uo = A[3 * n + NX(j)], c_uo = C[3 * n + NX(j)];
    int nb_ov = A[3 * n + PV(j)], c_ov = C[3 * n + PV(j)];
    int nb_pu = A[3 * t + PV(k)], c_pu = C[3 * t + PV(k)];
    int nb_vp = A[3 * t + NX(k)], c_vp = C[3 * t + NX(k)];
    set_tri(m, t, p, u, o);
    set_tri(m, n, p, o, v);
    A[3 * t + 0] = nb_uo; C[3 * t + 0] = (unsigned char)c_uo;
    A[3 * t + 1] = n;     C[3 * t + 1] = 0;
    A[3 * t + 2] = nb_pu; C[3 * t + 2] = (unsigned char)c_pu;
    A[3 * n + 0] = nb_ov; C[3 * n + 0] = (unsigned char)c_ov;
    A[3 * n + 1] = nb_vp; C[3 * n + 1] = (unsigned char)c_vp;
    A[3 * n + 2] = t;     C[3 * n + 2] = 0;
    if (nb_uo >= 0) A[3 * nb_uo + slot_of_nb(m, nb_uo, n)] = t;
    if (nb_vp >= 0) A[3 * nb_vp + slot_of_nb(m, nb_vp, t)] = n;
    m->flips++;
}

/* Lawson legalisation of the queued edges; (t,k) names the edge opposite the
 * vertex in slot k of t. */
static void legalize(mesh_t* m)
{
    while (m->stk_n > 0) {
        int e = m->stk[--m->stk_n];
        int t = e / 3, k = e % 3;
        int n = m->A[3 * t + k];
        if (n < 0 || m->C[3 * t + k]) continue;
        int j = slot_of_nb(m, n, t);
        int o = m->V[3 * n + j];
        if (INC(m, m->V[3 * t], m->V[3 * t + 1], m->V[3 * t + 2], o) > 0) {
            int p = m->V[3 * t + k];
            flip(m, t, k);
            /* p is now in slot 0 of both t and n */
            (void)p;
            push(m, t, 0);
            push(m, n, 0);
            if (m->general) {
                push(m, t, 2);
                push(m, n, 1);
            }
        }
    }
}

/* Visibility walk.  Returns the triangle containing p; *edge = -1 inside,
 * k if p lies on the edge opposite slot k, -2 if p coincides with a vertex. */
static int locate(const mesh_t* m, int p, int t, int* edge)
{
    const int* V = m->V;
    int start = 0;
    for (;;) {
        int moved = 0;
        for (int i = 0; i < 3; i++) {
            int k = (start + i) % 3;
            int u = V[3 * t + NX(k)], v = V[3 * t + PV(k)];
            if (ORI(m, u, v, p) < 0) {
                int n = m->A[3 * t + k];
                if (n < 0) { *edge = -3; return t; }
                t = n;
                start = (k + 1) % 3;
                moved = 1;
                break;
            }
        }
        if (!moved) break;
    }
    int zeros = 0, ke = -1;
    for (int k = 0; k < 3; k++) {
        if (ORI(m, V[3 * t + NX(k)], V[3 * t + PV(k)], p) == 0) { zeros++; ke = k; }
    }
    *edge = zeros == 0 ? -1 : (zeros == 1 ? ke : -2);
    return t;
}

static int insert_point(mesh_t* m, int p, int* hint)
{
    int edge;
    int t = locate(m, p, *hint, &edge);
    if (edge == -2) return OR_ERR_DUPLICATE;
    if (edge == -3) return OR_ERR_RANGE;
    int* V = m->V;
    int* A = m->A;
    unsigned char* C = m->C;
    if (edge == -1) {
        int a = V[3 * t], b = V[3 * t + 1], c = V[3 * t + 2];
        int na = A[3 * t], nb = A[3 * t + 1], nc = A[3 * t + 2];
        unsigned char ca = C[3 * t], cb = C[3 * t + 1], cc = C[3 * t + 2];
        int t1 = m->T, t2 = m->T + 1;
        m->T += 2;
        set_tri(m, t, p, b, c);
        set_tri(m, t1, p, c, a);
        set_tri(m, t2, p, a, b);
        A[3 * t] = na;  A[3 * t + 1] = t1;  A[3 * t + 2] = t2;
        A[3 * t1] = nb; A[3 * t1 + 1] = t2; A[3 * t1 + 2] = t;
        A[3 * t2] = nc; A[3 * t2 + 1] = t;  A[3 * t2 + 2] = t1;
        C[3 * t] = ca;  C[3 * t + 1] = 0;  C[3 * t + 2] = 0;
        C[3 * t1] = cb; C[3 * t1 + 1] = 0; C[3 * t1 + 2] = 0;
        C[3 * t2] = cc; C[3 * t2 + 1] = 0; C[3 * t2 + 2] = 0;
        if (nb >= 0) A[3 * nb + slot_of_nb(m, nb, t)] = t1;
        if (nc >= 0) A[3 * nc + slot_of_nb(m, nc, t)] = t2;
        push(m, t, 0);
        push(m, t1, 0);
        push(m, t2, 0);
        *hint = t;
    } else {
        int k = edge;
        int x = V[3 * t + k], u = V[3 * t + NX(k)], v = V[3 * t + PV(k)];
        int n = A[3 * t + k];
        if (n < 0) return OR_ERR_RANGE; /* on the square boundary */
        int j = slot_of_nb(m, n, t);
        int y = V[3 * n + j];
        int a_xu = A[3 * t + PV(k)], a_vx = A[3 * t + NX(k)];
        unsigned char c_xu = C[3 * t + PV(k)], c_vx = C[3 * t + NX(k)];
        int a_yv = A[3 * n + PV(j)], a_uy = A[3 * n + NX(j)];
        unsigned char c_yv = C[3 * n + PV(j)], c_uy = C[3 * n + NX(j)];
        unsigned char c_split = C[3 * t + k];
        int t2 = m->T, n2 = m->T + 1;
        m->T += 2;
        set_tri(m, t, x, u, p);
        set_tri(m, t2, x, p, v);
        set_tri(m, n, y, v, p);
        set_tri(m, n2, y, p, u);
        A[3 * t] = n2;    A[3 * t + 1] = t2;    A[3 * t + 2] = a_xu;
        A[3 * t2] = n;    A[3 * t2 + 1] = a_vx; A[3 * t2 + 2] = t;
        A[3 * n] = t2;    A[3 * n + 1] = n2;    A[3 * n + 2] = a_yv;
        A[3 * n2] = t;    A[3 * n2 + 1] = a_uy; A[3 * n2 + 2] = n;
        C[3 * t] = c_split;  C[3 * t + 1] = 0;     C[3 * t + 2] = c_xu;
        C[3 * t2] = c_split; C[3 * t2 + 1] = c_vx; C[3 * t2 + 2] = 0;
        C[3 * n] = c_split;  C[3 * n + 1] = 0;     C[3 * n + 2] = c_yv;
        C[3 * n2] = c_split; C[3 * n2 + 1] = c_uy; C[3 * n2 + 2] = 0;
        if (a_vx >= 0) A[3 * a_vx + slot_of_nb(m, a_vx, t)] = t2;
        if (a_uy >= 0) A[3 * a_uy + slot_of_nb(m, a_uy, n)] = n2;
        push(m, t, 2);
        push(m, t2, 1);
        push(m, n, 2);
        push(m, n2, 1);
        *hint = t;
    }
    legalize(m);
    return OR_OK;
}

/* ---------------------------------------------------------------- Hilbert order */

static uint64_t hilbert_d(uint32_t x, uint32_t y, int order)
{
    uint64_t d = 0;
    for (uint32_t s = 1u << (order - 1); s > 0; s >>= 1) {
        uint32_t rx = (x & s) > 0, ry = (y & s) > 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            uint32_t t = x; x = y; y = t;
        }
    }
    return d;
}

typedef struct { uint64_t key; int id; } kv_t;

static int cmp_kv(const void* a, const void* b)
{
    const kv_t* x = (const kv_t*)a;
    const kv_t* y = (const kv_t*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->id - y->id;
}

/* ---------------------------------------------------------------- segments */

typedef struct {
    int* tris; int nt, ct;       /* crossed triangles */
    int* L; int nl, cl;          /* left chain, a -> b */
    int* R; int nr, cr;          /* right chain, a -> b */
    int* bnd; int nb, cb;        /* boundary records: x, y, outer tri, cflag */
    kv_t* items; int ci;
    int* mark; int stamp;
} cav_t;

#define GROW(arr, n, c, type)                                                    \
    do {                                                                         \
        if ((n) >= (c)) {                                                        \
            int nc__ = (c) * 2 + 64;                                             \
            type* p__ = (type*)realloc((arr), sizeof(type) * (size_t)nc__);      \
            if (!p__) return OR_ERR_ALLOC;                                       \
            (arr) = p__;                                                         \
            (c) = nc__;                                                          \
        }                                                                        \
    } while (0)

static int cav_push_tri(cav_t* cv, int t) { GROW(cv->tris, cv->nt, cv->ct, int); cv->tris[cv->nt++] = t; return OR_OK; }
static int cav_push_L(cav_t* cv, int v) { if (cv->nl && cv->L[cv->nl - 1] == v) return OR_OK; GROW(cv->L, cv->nl, cv->cl, int); cv->L[cv->nl++] = v; return OR_OK; }
static int cav_push_R(cav_t* cv, int v) { if (cv->nr && cv->R[cv->nr - 1] == v) return OR_OK; GROW(cv->R, cv->nr, cv->cr, int); cv->R[cv->nr++] = v; return OR_OK; }

/* pseudo-polygon triangulation: base a->b, chain strictly left of a->b, in
 * boundary order from a to b.  Emits CCW triangles (a,b,c). */
static int tri_pseudo(const mesh_t* m, cav_t* cv, int a, int b, const int* ch, int n, int* out, int* no)
{
    if (n == 0) return OR_OK;
    int ci = 0;
    for (int i = 1; i < n; i++)
        if (INC(m, a, b, ch[ci], ch[i]) > 0) ci = i;
    int c = ch[ci];
    out[3 * (*no) + 0] = a;
    out[3 * (*no) + 1] = b;
    out[3 * (*no) + 2] = c;
    (*no)++;
    int r = tri_pseudo(m, cv, a, c, ch, ci, out, no);
    if (r) return r;
    return tri_pseudo(m, cv, c, b, ch + ci + 1, n - ci - 1, out, no);
}

static uint64_t ekey(int x, int y)
{
    uint32_t lo = (uint32_t)(x < y ? x : y), hi = (uint32_t)(x < y ? y : x);
    return ((uint64_t)lo << 32) | hi;
}

/* Retriangulate the cavity of segment (a,b) whose crossed triangles, chains and
 * boundary records were collected by the walk. */
static int retriangulate(mesh_t* m, cav_t* cv, int a, int b)
{
    int nt = cv->nt;
    if (cv->nl + cv->nr != nt) return OR_ERR_INTERNAL;
    int* verts = (int*)malloc(sizeof(int) * 3 * (size_t)nt);
    int* rev = (int*)malloc(sizeof(int) * (size_t)(cv->nr + 1));
    if (!verts || !rev) { free(verts); free(rev); return OR_ERR_ALLOC; }
    int no = 0;
    int r = tri_pseudo(m, cv, a, b, cv->L, cv->nl, verts, &no);
    for (int i = 0; i < cv->nr; i++) rev[i] = cv->R[cv->nr - 1 - i];
    if (!r) r = tri_pseudo(m, cv, b, a, rev, cv->nr, verts, &no);
    free(rev);
    if (r || no != nt) { free(verts); return r ? r : OR_ERR_INTERNAL; }
    for (int i = 0; i < nt; i++) {
        int t = cv->tris[i];
        set_tri(m, t, verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
        for (int k = 0; k < 3; k++) { m->A[3 * t + k] = -1; m->C[3 * t + k] = 0; }
    }
    free(verts);
    /* relink: half-edges of new triangles + outer sides of boundary records */
    int ni = 3 * nt + cv->nb;
    if (ni > cv->ci) {
        kv_t* it = (kv_t*)realloc(cv->items, sizeof(kv_t) * (size_t)ni);
        if (!it) return OR_ERR_ALLOC;
        cv->items = it;
        cv->ci = ni;
    }
    int q = 0;
    for (int i = 0; i < nt; i++) {
        int t = cv->tris[i];
        for (int k = 0; k < 3; k++) {
            cv->items[q].key = ekey(m->V[3 * t + NX(k)], m->V[3 * t + PV(k)]);
            cv->items[q].id = 3 * t + k; /* >= 0: new half-edge */
            q++;
        }
    }
    for (int i = 0; i < cv->nb; i++) {
        int* br = cv->bnd + 4 * i;
        cv->items[q].key = ekey(br[0], br[1]);
        cv->items[q].id = -1 - i; /* < 0: boundary record i */
        q++;
    }
    qsort(cv->items, (size_t)q, sizeof(kv_t), cmp_kv);
    int i = 0;
    while (i < q) {
        int j = i + 1;
        while (j < q && cv->items[j].key == cv->items[i].key) j++;
        int cnt = j - i;
        if (cnt == 2) {
            int e0 = cv->items[i].id, e1 = cv->items[i + 1].id;
            if (e0 >= 0 && e1 >= 0) {
                m->A[e0] = e1 / 3;
                m->A[e1] = e0 / 3;
                uint64_t kab = ekey(a, b);
                if (cv->items[i].key == kab) { m->C[e0] = 1; m->C[e1] = 1; }
            } else if (e0 >= 0 || e1 >= 0) {
                int e = e0 >= 0 ? e0 : e1;
                int bi = -1 - (e0 >= 0 ? e1 : e0);
                int* br = cv->bnd + 4 * bi;
                int ot = br[2];
                m->A[e] = ot;
                m->C[e] = (unsigned char)br[3];
                if (ot >= 0) {
                    /* the outer triangle's slot opposite the edge: the vertex not in it */
                    for (int k = 0; k < 3; k++) {
                        int w = m->V[3 * ot + k];
                        if (w != br[0] && w != br[1]) { m->A[3 * ot + k] = e / 3; break; }
                    }
                }
            } else {
                return OR_ERR_INTERNAL;
            }
        } else if (cnt == 1) {
            int e = cv->items[i].id;
            if (e < 0) return OR_ERR_INTERNAL;
            /* a boundary record whose outer side is the hull was dropped */
            m->A[e] = -1;
        } else {
            return OR_ERR_INTERNAL;
        }
        i = j;
    }
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++)
            if (!m->C[3 * cv->tris[t] + k]) push(m, cv->tris[t], k);
    m->general = 1;
    i64 before = m->flips;
    legalize(m);
    m->general = 0;
    m->cavity_flips += m->flips - before;
    return OR_OK;
}

static void mark_edge(mesh_t* m, int t, int k)
{
    m->C[3 * t + k] = 1;
    int n = m->A[3 * t + k];
    if (n >= 0) m->C[3 * n + slot_of_nb(m, n, t)] = 1;
}

static i64 dot3(const mesh_t* m, int a, int b, int c) /* (b-a).(c-a) */
{
    const int32_t* P = m->P;
    return (i64)(P[2 * b] - P[2 * a]) * (P[2 * c] - P[2 * a]) +
           (i64)(P[2 * b + 1] - P[2 * a + 1]) * (P[2 * c + 1] - P[2 * a + 1]);
}

/* Insert segment (a,b); splits it at vertices lying exactly on it. */
static int insert_segment(mesh_t* m, cav_t* cv, int a, int b, i64* nsplit)
{
    int guard = 0;
    while (a != b) {
        if (++guard > 4 * m->nv + 16) return OR_ERR_INTERNAL;
        /* rotate around a to find the wedge (a,u,w) containing direction a->b */
        int t0 = m->vt[a], t = t0;
        int found = 0, steps = 0;
        int k = -1;
        for (;;) {
            if (++steps > m->T + 4) return OR_ERR_INTERNAL;
            int s = 0;
            while (m->V[3 * t + s] != a) s++;
            int u = m->V[3 * t + NX(s)], w = m->V[3 * t + PV(s)];
            int ou = ORI(m, a, u, b);
            if (ou == 0 && dot3(m, a, u, b) > 0) {
                /* u lies on the ray a->b: edge (a,u) is a piece of the segment */
                mark_edge(m, t, PV(s));
                if (u != b) (*nsplit)++;
                a = u;
                found = 2;
                break;
            }
            int ow = ORI(m, a, w, b);
            if (ou > 0 && ow < 0) { k = s; found = 1; break; }
            /* rotate counter-clockwise around a: cross the edge (a,w), opposite u */
            int n = m->A[3 * t + NX(s)];
            /* input vertices are strictly inside the square: their star is closed */
            if (n < 0) return OR_ERR_INTERNAL;
            t = n;
            if (t == t0) return OR_ERR_INTERNAL;
        }
        if (found == 2) continue;
        /* walk across the crossing edges from triangle t (a in slot k) */
        cv->nt = cv->nl = cv->nr = cv->nb = 0;
        cv->stamp++;
        int u = m->V[3 * t + NX(k)], w = m->V[3 * t + PV(k)];
        /* u is right of a->b, w is left */
        int r = cav_push_tri(cv, t);
        if (r) return r;
        cav_push_R(cv, u);
        cav_push_L(cv, w);
        int cur = t, ce_r = u, ce_l = w; /* current crossing edge (right, left) */
        int q;
        for (;;) {
            int slot = -1;
            for (int s = 0; s < 3; s++) {
                int x = m->V[3 * cur + NX(s)], y = m->V[3 * cur + PV(s)];
                if ((x == ce_r && y == ce_l) || (x == ce_l && y == ce_r)) slot = s;
            }
            if (slot < 0) return OR_ERR_INTERNAL;
            if (m->C[3 * cur + slot]) return OR_ERR_CROSSING;
            int n = m->A[3 * cur + slot];
            if (n < 0) return OR_ERR_INTERNAL;
            int j = slot_of_nb(m, n, cur);
            q = m->V[3 * n + j];
            r = cav_push_tri(cv, n);
            if (r) return r;
            cur = n;
            if (q == b) break;
            int oq = ORI(m, a, b, q);
            if (oq == 0) { (*nsplit)++; break; }
            if (oq > 0) { ce_l = q; cav_push_L(cv, q); }
            else { ce_r = q; cav_push_R(cv, q); }
        }
        /* boundary records: edges of crossed triangles whose neighbour is not crossed */
        for (int i = 0; i < cv->nt; i++) cv->mark[cv->tris[i]] = cv->stamp;
        for (int i = 0; i < cv->nt; i++) {
            int tt = cv->tris[i];
            for (int s = 0; s < 3; s++) {
                int n = m->A[3 * tt + s];
                if (n >= 0 && cv->mark[n] == cv->stamp) continue;
                if (4 * cv->nb + 4 > cv->cb) {
                    int nc2 = cv->cb * 2 + 256;
                    int* nbuf = (int*)realloc(cv->bnd, sizeof(int) * (size_t)nc2);
                    if (!nbuf) return OR_ERR_ALLOC;
                    cv->bnd = nbuf;
                    cv->cb = nc2;
                }
                int* br = cv->bnd + 4 * cv->nb;
                br[0] = m->V[3 * tt + NX(s)];
                br[1] = m->V[3 * tt + PV(s)];
                br[2] = n;
                br[3] = m->C[3 * tt + s];
                cv->nb++;
            }
        }
        r = retriangulate(m, cv, a, q);
        if (r) return r;
        a = q;
    }
    return OR_OK;
}

/* ---------------------------------------------------------------- driver */

/*
 * oracle_cdt: pts = 2*n int32 coordinates on the 2^30 grid, strictly inside
 * (0, 2^30-1); cons = 2*nc vertex ids.  Writes 3*T vertex ids into tris
 * (capacity 3*(2n+2)), corners are ids n..n+3.  stats (may be NULL, 4 i64):
 * flips, segment splits, T, flips after cavity retriangulation.
 */
int oracle_cdt(const int32_t* pts, int n, const int32_t* cons, int nc, int32_t* tris, int* out_t,
               i64* stats)
{
    int nv = n + 4;
    int32_t* P = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)nv);
    mesh_t m;
    memset(&m, 0, sizeof(m));
    cav_t cv;
    memset(&cv, 0, sizeof(cv));
    int rc = OR_OK;
    kv_t* order = NULL;
    if (!P) return OR_ERR_ALLOC;
    memcpy(P, pts, sizeof(int32_t) * 2 * (size_t)n);
    for (int i = 0; i < n; i++) {
        if (P[2 * i] <= 0 || P[2 * i] >= GRID_R || P[2 * i + 1] <= 0 || P[2 * i + 1] >= GRID_R) {
            free(P);
            return OR_ERR_RANGE;
        }
    }
    P[2 * n] = 0;          P[2 * n + 1] = 0;
    P[2 * n + 2] = GRID_R; P[2 * n + 3] = 0;
    P[2 * n + 4] = GRID_R; P[2 * n + 5] = GRID_R;
    P[2 * n + 6] = 0;      P[2 * n + 7] = GRID_R;
    m.P = P;
    m.nv = nv;
    m.cap = 2 * n + 2 > 2 ? 2 * n + 2 : 2;
    m.V = (int*)malloc(sizeof(int) * 3 * (size_t)m.cap);
    m.A = (int*)malloc(sizeof(int) * 3 * (size_t)m.cap);
    m.C = (unsigned char*)calloc(3 * (size_t)m.cap, 1);
    m.vt = (int*)malloc(sizeof(int) * (size_t)nv);
    cv.mark = (int*)calloc((size_t)m.cap, sizeof(int));
    order = (kv_t*)malloc(sizeof(kv_t) * (size_t)(n > 0 ? n : 1));
    if (!m.V || !m.A || !m.C || !m.vt || !cv.mark || !order) { rc = OR_ERR_ALLOC; goto done; }
    set_tri(&m, 0, n, n + 1, n + 3);
    set_tri(&m, 1, n + 1, n + 2, n + 3);
    m.A[0] = 1;  m.A[1] = -1; m.A[2] = -1;
    m.A[3] = -1; m.A[4] = 0;  m.A[5] = -1;
    m.T = 2;
    for (int i = 0; i < n; i++) {
        order[i].key = hilbert_d((uint32_t)P[2 * i] >> 14, (uint32_t)P[2 * i + 1] >> 14, 16);
        order[i].id = i;
    }
    qsort(order, (size_t)n, sizeof(kv_t), cmp_kv);
    int hint = 0;
    for (int i = 0; i < n; i++) {
        rc = insert_point(&m, order[i].id, &hint);
        if (rc) goto done;
    }
    i64 nsplit = 0;
    for (int i = 0; i < nc; i++) {
        int a = cons[2 * i], b = cons[2 * i + 1];
        if (a < 0 || a >= n || b < 0 || b >= n || a == b) { rc = OR_ERR_CONSTRAINT; goto done; }
        rc = insert_segment(&m, &cv, a, b, &nsplit);
        if (rc) goto done;
    }
    memcpy(tris, m.V, sizeof(int32_t) * 3 * (size_t)m.T);
    *out_t = m.T;
    if (stats) {
        stats[0] = m.flips;
        stats[1] = nsplit;
        stats[2] = m.T;
        stats[3] = m.cavity_flips;
    }
done:
    free(P); free(m.V); free(m.A); free(m.C); free(m.vt); free(m.stk);
    free(cv.tris); free(cv.L); free(cv.R); free(cv.bnd); free(cv.items); free(cv.mark);
    free(order);
    return rc;
}

/* ---------------------------------------------------------------- validator */

typedef struct { uint64_t key; int he; } hk_t;

static int cmp_hk(const void* a, const void* b)
{
    const hk_t* x = (const hk_t*)a;
    const hk_t* y = (const hk_t*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->he - y->he;
}

static int find_edge(const hk_t* hs, int nh, uint64_t key)
{
    int lo = 0, hi = nh;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (hs[mid].key < key) lo = mid + 1; else hi = mid;
    }
    return (lo < nh && hs[lo].key == key) ? lo : -1;
}

/*
 * oracle_validate: exact check of a triangulation of the bounding square.
 *   pts: 2*n coordinates (corners are appended internally at ids n..n+3)
 *   tris: 3*nt vertex ids; cons: 2*nc vertex-id pairs
 * out[12] (i64):
 *   0 triangles not strictly CCW        1 vertex ids out of range
 *   2 area mismatch (0/1)               3 non-manifold / unmatched edges
 *   4 boundary edges that are not square sides   5 vertices never used
 *   6 constraints whose edge chain is missing    7 unconstrained edges with
 *     inCircle > 0 (not locally Delaunay)        8 unconstrained interior edges
 *     with inCircle == 0 (cocircular; the triangulation is not unique)
 *   9 triangle count != 2(n+4)-6        10 constraint edges (pieces) found
 *   11 interior edges checked
 * Returns 0 when the triangulation is a valid CDT (items 0-7 and 9 zero).
 */
int oracle_validate(const int32_t* pts, int n, const int32_t* tris, int nt, const int32_t* cons,
                    int nc, i64* out)
{
    int nv = n + 4;
    for (int i = 0; i < 12; i++) out[i] = 0;
    int32_t* P = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)nv);
    hk_t* hs = (hk_t*)malloc(sizeof(hk_t) * 3 * (size_t)(nt > 0 ? nt : 1));
    unsigned char* used = (unsigned char*)calloc((size_t)nv, 1);
    unsigned char* cmark = (unsigned char*)calloc(3 * (size_t)(nt > 0 ? nt : 1), 1);
    int* deg = (int*)calloc((size_t)nv + 1, sizeof(int));
    int* adjv = NULL;
    int rc = 0;
    if (!P || !hs || !used || !cmark || !deg) { rc = -1; goto vdone; }
    memcpy(P, pts, sizeof(int32_t) * 2 * (size_t)n);
    P[2 * n] = 0;          P[2 * n + 1] = 0;
    P[2 * n + 2] = GRID_R; P[2 * n + 3] = 0;
    P[2 * n + 4] = GRID_R; P[2 * n + 5] = GRID_R;
    P[2 * n + 6] = 0;      P[2 * n + 7] = GRID_R;
    mesh_t m;
    memset(&m, 0, sizeof(m));
    m.P = P;
    m.nv = nv;
    if (nt != 2 * nv - 6) out[9] = 1;
    i128 area = 0;
    for (int t = 0; t < nt; t++) {
        int a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        if (a < 0 || a >= nv || b < 0 || b >= nv || c < 0 || c >= nv) { out[1]++; continue; }
        used[a] = used[b] = used[c] = 1;
        if (ORI(&m, a, b, c) <= 0) out[0]++;
        area += (i128)((i64)P[2 * a] * P[2 * b + 1] - (i64)P[2 * b] * P[2 * a + 1]);
        area += (i128)((i64)P[2 * b] * P[2 * c + 1] - (i64)P[2 * c] * P[2 * b + 1]);
        area += (i128)((i64)P[2 * c] * P[2 * a + 1] - (i64)P[2 * a] * P[2 * c + 1]);
    }
    if (out[1]) { rc = 1; goto vdone; }
    if (area != (i128)2 * GRID_R * (i128)GRID_R) out[2] = 1;
    for (int v = 0; v < nv; v++) if (!used[v]) out[5]++;
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) {
            hs[3 * t + k].key = ekey(tris[3 * t + NX(k)], tris[3 * t + PV(k)]);
            hs[3 * t + k].he = 3 * t + k;
        }
    int nh = 3 * nt;
    qsort(hs, (size_t)nh, sizeof(hk_t), cmp_hk);
    /* adjacency lists for the constraint-chain walk */
    for (int i = 0; i < nh; i++) {
        uint64_t key = hs[i].key;
        if (i + 1 < nh && hs[i + 1].key == key) continue; /* count each undirected edge once */
        deg[(int)(key >> 32)]++;
        deg[(int)(key & 0xffffffffu)]++;
    }
    {
        int s = 0;
        for (int v = 0; v <= nv; v++) { int d = deg[v]; deg[v] = s; s += d; }
        adjv = (int*)malloc(sizeof(int) * (size_t)(s > 0 ? s : 1));
        if (!adjv) { rc = -1; goto vdone; }
        int* fill = (int*)malloc(sizeof(int) * (size_t)nv);
        if (!fill) { rc = -1; goto vdone; }
        for (int v = 0; v < nv; v++) fill[v] = deg[v];
        for (int i = 0; i < nh; i++) {
            uint64_t key = hs[i].key;
            if (i + 1 < nh && hs[i + 1].key == key) continue;
            int x = (int)(key >> 32), y = (int)(key & 0xffffffffu);
            adjv[fill[x]++] = y;
            adjv[fill[y]++] = x;
        }
        free(fill);
    }
    /* constraint chains */
    for (int c = 0; c < nc; c++) {
        int a = cons[2 * c], b = cons[2 * c + 1];
        if (a < 0 || a >= nv || b < 0 || b >= nv || a == b) { out[6]++; continue; }
        int cur = a, ok = 1, guard = 0;
        while (cur != b) {
            if (++guard > nv) { ok = 0; break; }
            int nxt = -1;
            for (int e = deg[cur]; e < deg[cur + 1]; e++) {
                int v = adjv[e];
                if (v == b) { nxt = b; break; }
                if (ORI(&m, a, b, v) != 0) continue;
                /* strictly between cur and b along the segment */
                if (dot3(&m, cur, v, b) > 0 && dot3(&m, b, v, cur) > 0) { nxt = v; break; }
            }
            if (nxt < 0) { ok = 0; break; }
            int pos = find_edge(hs, nh, ekey(cur, nxt));
            while (pos >= 0 && pos < nh && hs[pos].key == ekey(cur, nxt)) { cmark[hs[pos].he] = 1; pos++; }
            out[10]++;
            cur = nxt;
        }
        if (!ok) out[6]++;
    }
    /* edge matching + local Delaunay test */
    int i = 0;
    while (i < nh) {
        int j = i + 1;
        while (j < nh && hs[j].key == hs[i].key) j++;
        int cnt = j - i;
        if (cnt == 1) {
            int x = (int)(hs[i].key >> 32), y = (int)(hs[i].key & 0xffffffffu);
            int side = (x == n && y == n + 1) || (x == n + 1 && y == n + 2) ||
                       (x == n + 2 && y == n + 3) || (x == n && y == n + 3);
            if (!side) out[4]++;
        } else if (cnt == 2) {
            int h0 = hs[i].he, h1 = hs[i + 1].he;
            int t0 = h0 / 3, k0 = h0 % 3, t1 = h1 / 3, k1 = h1 % 3;
            /* opposite directions */
            if (tris[3 * t0 + NX(k0)] != tris[3 * t1 + PV(k1)] || tris[3 * t0 + PV(k0)] != tris[3 * t1 + NX(k1)]) {
                out[3]++;
            } else if (!cmark[h0] && !cmark[h1]) {
                int o = tris[3 * t1 + k1];
                int s = INC(&m, tris[3 * t0], tris[3 * t0 + 1], tris[3 * t0 + 2], o);
                out[11]++;
                if (s > 0) out[7]++;
                else if (s == 0) out[8]++;
            }
        } else {
            out[3]++;
        }
        i = j;
    }
    rc = (out[0] || out[2] || out[3] || out[4] || out[5] || out[6] || out[7] || out[9]) ? 1 : 0;
vdone:
    free(P); free(hs); free(used); free(cmark); free(deg); free(adjv);
    return rc;
}
