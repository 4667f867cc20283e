// Constraint-segment recovery on gfx950.
//
// Reference: addConstraints (addConstraints.cu:1723-1835) locates each
// segment, collects the triangles it crosses, builds the upper/lower vertex
// chains and retriangulates each chain (formPolygon_s1/s2, calDistance,
// makeTree, reTriangle, :1491-1721).  It drops segments that pass through a
// vertex (:962-965).  Here:
//   1. split pass: every segment is walked through the Delaunay mesh and cut
//      at the vertices lying exactly on it into pieces (a, b) with no vertex
//      in their open interior; pieces that already are edges are marked
//      constrained at once;
//   2. rounds: every pending piece walks its crossed triangles and claims
//      them (atomicMin of a per-round key, smallest piece id wins); a piece
//      that owns all its crossed triangles retriangulates its cavity in
//      place: the left and right pseudo-polygons are triangulated by the
//      recursive Delaunay-vertex rule (for base (A,B) pick the chain vertex
//      C whose circle through A,B contains no other chain vertex, recurse on
//      (A,C) and (C,B)), which reuses exactly the crossed slots (k crossing
//      edges -> k+1 triangles before and after);
//   3. outer adjacency is resolved after all cavities of the round are
//      written (a neighbour rewritten in the same round is found in its
//      cavity's slot list);
// and the caller finishes with constrained flips.
#include <algorithm>
#include <climits>

#include "cons_bodies.hpp"
#include "mesh.hpp"

namespace gcdt {

__global__ __launch_bounds__(kBlock) void k_split_count(Cons c, const int2* cons, int C, int nmax, int* cnt)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < C; i += (long long)gridDim.x * blockDim.x) {
        int2 e = cons[i];
        if (e.x < 0 || e.y < 0 || e.x >= nmax || e.y >= nmax || e.x == e.y) {
            atomicOr(c.err, MERR_INVALID);
            cnt[i] = 0;
            continue;
        }
        cnt[i] = split_walk(c, e.x, e.y, nullptr);
    }
}

__global__ __launch_bounds__(kBlock) void k_split_emit(Cons c, const int2* cons, int C, const int* off, int2* pieces)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < C; i += (long long)gridDim.x * blockDim.x) {
        int2 e = cons[i];
        if (off[i + 1] == off[i]) continue;
        split_walk(c, e.x, e.y, pieces + off[i]);
    }
}

__global__ __launch_bounds__(kBlock) void k_build_vt(const int3* tri, int T, int* vt)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        int3 t = tri[i];
        vt[t.x] = (int)i;
        vt[t.y] = (int)i;
        vt[t.z] = (int)i;
    }
}

// queue of pieces that need a cavity (first vertex >= 0)
__global__ __launch_bounds__(kBlock) void k_piece_queue(const int2* pieces, int np, int* q, int* nq)
{
    WAVE_LOOP(i, np)
    {
        bool want = i < np && pieces[i].x >= 0;
        int slot = wave_alloc(want ? 1 : 0, nq);
        if (want) q[slot] = (int)i;
    }
}

// ---------------------------------------------------------------- rounds

// R1: claim crossed triangles; pieces that became edges are retired.
__global__ __launch_bounds__(kBlock) void k_cav_claim(Cons c, const int2* pieces, const int* q, const int* nq,
                                                      unsigned epoch, int* ncross, int* total)
{
    int n = *nq;
    WAVE_LOOP(i, n)
    {
        int k = 0;
        if (i < n) {
            int pi = q[i];
            int2 pc = pieces[pi];
            unsigned long long key = cav_key(epoch, pi);
            k = cav_walk(c, pc.x, pc.y, true, nullptr, [&](int t) { atomicMin(&c.owner[t], key); });
            if (k < 0) {
                atomicOr(c.err, MERR_WALK);
                k = 0;
            }
            ncross[i] = k;
        }
        int tot;
        wave_excl_scan(k, tot);
        if (__lane_id() == 0 && tot) atomicAdd(total, tot);
    }
}

__global__ __launch_bounds__(kBlock) void k_cav_check(Cons c, const int2* pieces, const int* q, const int* nq,
                                                      const int* ncross, unsigned epoch, int* q2, int* nq2, Win* wins,
                                                      int* nwin, int* scr_cursor, long long scr_cap)
{
    int n = *nq;
    WAVE_LOOP(i, n)
    {
        int pi = -1, K = 0, t0 = -1;
        bool win = false, keep = false;
        if (i < n) {
            pi = q[i];
            K = ncross[i];
            if (K > 0) {
                int2 pc = pieces[pi];
                unsigned long long key = cav_key(epoch, pi);
                bool all = true;
                cav_walk(c, pc.x, pc.y, false, &t0, [&](int t) { all &= c.owner[t] == key; });
                win = all;
                keep = !all;
            }
        }
        int need = win ? (int)cav_scratch_ints(K) : 0;
        int off = wave_alloc(need, scr_cursor);
        if (win && off + (long long)need > scr_cap) {
            win = false;  // defer to a later round
            keep = true;
        }
        int ws = wave_alloc(win ? 1 : 0, nwin);
        if (win) wins[ws] = Win{pi, t0, K, off};
        int qs = wave_alloc(keep ? 1 : 0, nq2);
        if (keep) q2[qs] = pi;
    }
}

// R3: collect the cavity and retriangulate it in place.
__global__ __launch_bounds__(kBlock) void k_cav_retri(Cons c, const int2* pieces, const Win* wins, const int* nwin,
                                                      int* scr, unsigned epoch, int* rew, int* nrew)
{
    int nw = *nwin;
    for (long long wi = (long long)blockIdx.x * blockDim.x + threadIdx.x; wi < nw; wi += (long long)gridDim.x * blockDim.x) {
        if (!cav_retri_one(c, pieces, wins, (int)wi, scr, epoch)) continue;
        Win w = wins[wi];
        int base = atomicAdd(nrew, w.K);
        for (int j = 0; j < w.K; j++) rew[base + j] = scr[w.off + j];
    }
}

// R4: resolve the marked outer adjacency of rewritten triangles
__global__ __launch_bounds__(kBlock) void k_cav_fix(Cons c, const Win* wins, const int* scr, const int* rew,
                                                    const int* nrew, unsigned epoch)
{
    int n = *nrew;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        cav_fix_one(c, wins, scr, rew[i], epoch);
}

__global__ void k_cav_begin(int* cnt)
{
    cnt[C_NQ2] = 0;
    cnt[C_NWIN] = 0;
    cnt[C_SCR] = 0;
    cnt[C_NREW] = 0;
    cnt[C_TOTAL] = 0;
}

// Split pass: pieces of every constraint (existing edges marked constrained
// and stored with a negative first vertex).  Returns the piece count.
static int split_pieces(Ctx& cx, const Cons& c, const int2* cons, int C, int nmax, DBuf<int2>& pieces)
{
    DBuf<int> cnt((size_t)C + 1);
    k_split_count<<<grid_for(C), kBlock, 0, cx.st>>>(c, cons, C, nmax, cnt.p);
    GCDT_LAUNCH_CHECK();
    cx.check_err();
    scan_excl_total(cx, cnt.p, cnt.p, C);
    int np = cx.get(C_TOTAL);
    pieces.alloc(np > 0 ? np : 1);
    k_split_emit<<<grid_for(C), kBlock, 0, cx.st>>>(c, cons, C, cnt.p, pieces.p);
    GCDT_LAUNCH_CHECK();
    cx.check_err();
    return np;
}

long long mark_constraint_chains(Ctx& cx, const Mesh& m, int T, int nv, int nmax, const int2* cons, int C)
{
    if (C <= 0) return 0;
    DBuf<int> vt(nv);
    fill_i32(cx, vt.p, -1, nv);
    k_build_vt<<<grid_for(T), kBlock, 0, cx.st>>>(m.tri, T, vt.p);
    GCDT_LAUNCH_CHECK();
    Cons c;
    c.m = m;
    c.vt = vt.p;
    c.owner = nullptr;
    c.stamp = nullptr;
    c.cavof = nullptr;
    c.err = cx.dcnt.p + C_ERR;
    DBuf<int2> pieces;
    return split_pieces(cx, c, cons, C, nmax, pieces);
}

__global__ __launch_bounds__(kBlock) void k_touched(const int* stamp, int T, int* touched)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x)
        touched[i] = stamp[i] != 0;
}

ConsResult insert_constraints(Ctx& cx, const Mesh& m, int T, int nv, int nmax, const int2* cons, int C,
                              int* touched, int* ntouched)
{
    (void)ntouched;
    ConsResult res;
    if (C <= 0) return res;
    DBuf<int> vt(nv);
    DBuf<unsigned long long> owner(T);
    DBuf<int> stamp(T), cavof(T);
    fill_i32(cx, vt.p, -1, nv);
    fill_u64(cx, owner.p, ~0ull, T);
    fill_i32(cx, stamp.p, 0, T);
    k_build_vt<<<grid_for(T), kBlock, 0, cx.st>>>(m.tri, T, vt.p);
    GCDT_LAUNCH_CHECK();
    Cons c;
    c.m = m;
    c.vt = vt.p;
    c.owner = owner.p;
    c.stamp = stamp.p;
    c.cavof = cavof.p;
    c.err = cx.dcnt.p + C_ERR;
    // 1. split into pieces
    DBuf<int2> pieces;
    int np = split_pieces(cx, c, cons, C, nmax, pieces);
    res.pieces = np;
    // 2. rounds over pieces that need a cavity
    DBuf<int> qa(np > 0 ? np : 1), qb(np > 0 ? np : 1), ncross(np > 0 ? np : 1);
    DBuf<Win> wins(np > 0 ? np : 1);
    cx.set(C_NQ, 0);
    k_piece_queue<<<grid_for(np), kBlock, 0, cx.st>>>(pieces.p, np, qa.p, cx.dcnt.p + C_NQ);
    GCDT_LAUNCH_CHECK();
    int nq = cx.get(C_NQ);
    int* Q = qa.p;
    int* Q2 = qb.p;
    DBuf<int> scr, rew;
    int r = 0;
    while (nq > 0) {
        if (++r > 100000) throw Error(GCDT_ERR_NOT_CONVERGED, "constraint rounds did not converge");
        unsigned ep = cx.next_epoch();
        k_cav_begin<<<1, 1, 0, cx.st>>>(cx.dcnt.p);
        GCDT_LAUNCH_CHECK();
        int g = grid_for(nq);
        k_cav_claim<<<g, kBlock, 0, cx.st>>>(c, pieces.p, Q, cx.dcnt.p + C_NQ, ep, ncross.p, cx.dcnt.p + C_TOTAL);
        GCDT_LAUNCH_CHECK();
        long long tot = cx.get(C_TOTAL);
        long long cap = 11LL * tot + 16LL * nq + 1024;
        if (cap > (1LL << 30)) cap = 1LL << 30;
        scr.ensure((size_t)cap);
        rew.ensure((size_t)tot + 1);
        k_cav_check<<<g, kBlock, 0, cx.st>>>(c, pieces.p, Q, cx.dcnt.p + C_NQ, ncross.p, ep, Q2, cx.dcnt.p + C_NQ2,
                                             wins.p, cx.dcnt.p + C_NWIN, cx.dcnt.p + C_SCR, (long long)scr.n);
        GCDT_LAUNCH_CHECK();
        k_cav_retri<<<g, kBlock, 0, cx.st>>>(c, pieces.p, wins.p, cx.dcnt.p + C_NWIN, scr.p, ep, rew.p,
                                             cx.dcnt.p + C_NREW);
        GCDT_LAUNCH_CHECK();
        k_cav_fix<<<grid_for(tot), kBlock, 0, cx.st>>>(c, wins.p, scr.p, rew.p, cx.dcnt.p + C_NREW, ep);
        GCDT_LAUNCH_CHECK();
        cx.read(0, 16);
        if (cx.hcnt[C_ERR]) cx.check_err();
        res.rewritten += cx.hcnt[C_NREW];
        nq = cx.hcnt[C_NQ2];
        if (nq > 0 && cx.hcnt[C_NWIN] == 0) throw Error(GCDT_ERR_INTERNAL, "constraint round without progress");
        cx.set(C_NQ, nq);
        std::swap(Q, Q2);
    }
    res.rounds = r;
    if (touched) {
        k_touched<<<grid_for(T), kBlock, 0, cx.st>>>(stamp.p, T, touched);
        GCDT_LAUNCH_CHECK();
    }
    return res;
}

}  // namespace gcdt
