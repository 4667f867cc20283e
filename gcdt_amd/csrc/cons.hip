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
            // steiner mode: a piece that crosses a recovered one is deferred
            // before it claims anything (its claims would block others)
            if (c.steiner && cav_walk(c, pc.x, pc.y, false, nullptr, [](int) {}) == -2) {
                c.defer[pi] = 1;
                k = 0;
            } else {
                k = cav_walk(c, pc.x, pc.y, true, nullptr, [&](int t) { atomicMin(&c.owner[t], key); });
                if (k < 0) {
                    atomicOr(c.err, MERR_WALK);
                    k = 0;
                }
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

// Cavities of more crossed triangles than this are retriangulated by a whole
// wave (k_cav_retri_wave): a long piece's pseudo-polygons have hundreds of
// vertices, and one thread's Delaunay-vertex scans over them took most of a
// solve with long segments.
#ifndef GCDT_CAV_WAVE_K
#define GCDT_CAV_WAVE_K 64
#endif
constexpr int CAV_WAVE_K = GCDT_CAV_WAVE_K;

// R3: collect the cavity and retriangulate it in place (one thread per cavity
// of up to CAV_WAVE_K triangles).
__global__ __launch_bounds__(kBlock) void k_cav_retri(Cons c, const int2* pieces, const Win* wins, const int* nwin,
                                                      int* scr, unsigned epoch, int* rew, int* nrew)
{
    int nw = *nwin;
    for (long long wi = (long long)blockIdx.x * blockDim.x + threadIdx.x; wi < nw; wi += (long long)gridDim.x * blockDim.x) {
        Win w = wins[wi];
        if (w.K > CAV_WAVE_K) continue;
        if (!cav_retri_one(c, pieces, wins, (int)wi, scr, epoch)) continue;
        int base = atomicAdd(nrew, w.K);
        for (int j = 0; j < w.K; j++) rew[base + j] = scr[w.off + j];
    }
}

// Delaunay vertex found by the 64 lanes of a wave: each scans every 64th
// candidate as DvScan does, then a butterfly keeps the better of two
// candidates -- the one strictly inside the other's circle, the lower index
// when cocircular -- which is the vertex DvScan returns.
struct DvWave {
    __device__ static int find(const int* PX, const int* PY, int lo, int hi, int2 PA, int2 PB)
    {
        const int lane = __lane_id();
        int ci = -1;
        int2 PC = make_int2(0, 0);
        for (int i = lo + lane; i < hi; i += 64) {
            const int2 PI = make_int2(PX[i], PY[i]);
            if (ci < 0 || incircle(PA, PB, PC, PI) > 0) {
                ci = i;
                PC = PI;
            }
        }
        for (int d = 32; d > 0; d >>= 1) {
            const int oi = __shfl(ci, lane ^ d, 64);
            const int ox = __shfl(PC.x, lane ^ d, 64), oy = __shfl(PC.y, lane ^ d, 64);
            if (oi < 0) continue;
            const int2 PO = make_int2(ox, oy);
            bool take = ci < 0;
            if (!take) {
                const int r = incircle(PA, PB, PC, PO);
                take = r > 0 || (r == 0 && oi < ci);
            }
            if (take) {
                ci = oi;
                PC = PO;
            }
        }
        return ci;
    }
};

// R3 for cavities above CAV_WAVE_K: every lane of the wave runs the body for
// the same cavity (the same loads and identical stores, so each lane reads
// its own writes) and the lanes share the Delaunay-vertex scans.
__global__ __launch_bounds__(kBlock) void k_cav_retri_wave(Cons c, const int2* pieces, const Win* wins,
                                                           const int* nwin, int* scr, unsigned epoch, int* rew,
                                                           int* nrew)
{
    const int nw = *nwin;
    const int lane = __lane_id();
    const long long wave0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long wi = wave0; wi < nw; wi += nwaves) {
        const Win w = wins[wi];
        if (w.K <= CAV_WAVE_K) continue;
        if (!cav_retri_one<DvWave>(c, pieces, wins, (int)wi, scr, epoch)) continue;
        int base = 0;
        if (lane == 0) base = atomicAdd(nrew, w.K);
        base = __shfl(base, 0, 64);
        for (int j = lane; j < w.K; j += 64) rew[base + j] = scr[w.off + j];
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

// ---------------------------------------------------------------- split by cells
//
// The split pass above walks each segment on one thread, vertex by vertex: a
// segment that runs through many collinear vertices (the lattice's diagonals:
// about 190 each) is a long chain of dependent loads on a single thread.  When
// the points are in Morton order of their 1024^2 cells (the single solve), the
// vertices on a segment are found instead by testing the points of every cell
// the segment touches, one (segment, cell) pair per thread; each segment's
// vertices are then sorted along it in LDS, and each piece between two
// consecutive ones is checked for being an existing edge on a thread of its
// own.  The pieces are the same as the walk's: in order along the segment,
// existing edges marked constrained and stored with a negative first vertex.
// Segments with more than SEG_SORT_MAX vertices on them keep the walk.

constexpr int SEG_SORT_MAX = 1024;

__host__ __device__ inline long long floor_div_ll(long long a, long long b)  // b > 0
{
    long long q = a / b;
    return (a % b != 0 && a < 0) ? q - 1 : q;
}

// Cells of column cx that the closed segment AB (A.x <= B.x) can touch: rows
// [lo, hi] of the y-range the segment spans inside the column (a superset).
__host__ __device__ inline void seg_column(int2 A, int2 B, int cx, int& lo, int& hi)
{
    if (A.x == B.x) {
        lo = min(A.y, B.y) >> 20;
        hi = max(A.y, B.y) >> 20;
        return;
    }
    long long dx = (long long)B.x - A.x, dy = (long long)B.y - A.y;
    long long xl = max((long long)A.x, (long long)cx << 20), xr = min((long long)B.x, ((long long)(cx + 1) << 20) - 1);
    long long nl = (long long)A.y * dx + (xl - A.x) * dy, nr = (long long)A.y * dx + (xr - A.x) * dy;
    long long ymin = floor_div_ll(min(nl, nr), dx), ymax = -floor_div_ll(-max(nl, nr), dx);
    lo = (int)max(0LL, min(1023LL, ymin >> 20));
    hi = (int)max(0LL, min(1023LL, ymax >> 20));
}

__device__ inline bool seg_valid(int2 e, int nmax) { return e.x >= 0 && e.y >= 0 && e.x < nmax && e.y < nmax && e.x != e.y; }

// endpoints ordered by x (the column walk's direction)
__device__ inline void seg_ends(const int2* pts, int2 e, int2& A, int2& B)
{
    A = pts[e.x];
    B = pts[e.y];
    if (A.x > B.x) {
        int2 t = A;
        A = B;
        B = t;
    }
}

__device__ inline unsigned cell_morton(unsigned x, unsigned y)
{
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    y = (y | (y << 8)) & 0x00FF00FFu;
    y = (y | (y << 4)) & 0x0F0F0F0Fu;
    y = (y | (y << 2)) & 0x33333333u;
    y = (y | (y << 1)) & 0x55555555u;
    return x | (y << 1);
}

__global__ __launch_bounds__(kBlock) void k_seg_ncells(const int2* pts, const int2* cons, int C, int nmax, int* cnt,
                                                       int* err)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < C; i += (long long)gridDim.x * blockDim.x) {
        int2 e = cons[i];
        if (!seg_valid(e, nmax)) {
            atomicOr(err, MERR_INVALID);
            cnt[i] = 0;
            continue;
        }
        int2 A, B;
        seg_ends(pts, e, A, B);
        int k = 0;
        for (int cx = A.x >> 20; cx <= (B.x >> 20); cx++) {
            int lo, hi;
            seg_column(A, B, cx, lo, hi);
            k += hi - lo + 1;
        }
        cnt[i] = k;
    }
}

// one (segment, Morton cell) pair per entry
__global__ __launch_bounds__(kBlock) void k_seg_cells(const int2* pts, const int2* cons, int C, int nmax, const int* off,
                                                      int* pair_seg, int* pair_cell)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < C; i += (long long)gridDim.x * blockDim.x) {
        int2 e = cons[i];
        if (!seg_valid(e, nmax)) continue;
        int2 A, B;
        seg_ends(pts, e, A, B);
        int o = off[i];
        for (int cx = A.x >> 20; cx <= (B.x >> 20); cx++) {
            int lo, hi;
            seg_column(A, B, cx, lo, hi);
            for (int cy = lo; cy <= hi; cy++) {
                pair_seg[o] = (int)i;
                pair_cell[o] = (int)cell_morton((unsigned)cx, (unsigned)cy);
                o++;
            }
        }
    }
}

// The points of each pair's cell that lie strictly inside its segment: counted
// per segment (pos == nullptr), or written to hits[hoff[seg] + ...].
__global__ __launch_bounds__(kBlock) void k_seg_hits(const int2* pts, const int* cell_end, const int2* cons,
                                                     const int* pair_seg, const int* pair_cell, long long npairs,
                                                     int* hcnt, const int* hoff, int* hits)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < npairs; i += (long long)gridDim.x * blockDim.x) {
        int sg = pair_seg[i], cell = pair_cell[i];
        int2 e = cons[sg];
        int2 A = pts[e.x], B = pts[e.y];
        int p0 = cell ? cell_end[cell - 1] : 0, p1 = cell_end[cell];
        for (int p = p0; p < p1; p++) {
            if (p == e.x || p == e.y) continue;
            int2 P = pts[p];
            if (orient2d(A, B, P) != 0 || !strictly_between(A, B, P)) continue;
            int k = atomicAdd(&hcnt[sg], 1);
            if (hits) hits[hoff[sg] + k] = p;
        }
    }
}

// pieces per segment: h + 1 for a sorted segment; a segment with more than
// SEG_SORT_MAX vertices on it goes to the walk list (its count comes later)
__global__ __launch_bounds__(kBlock) void k_seg_npieces(const int2* cons, int C, int nmax, const int* hcnt, int* np,
                                                        int* sort_list, int* nsort, int* walk_list, int* nwalk)
{
    WAVE_LOOP(i, C)
    {
        int h = 0;
        bool valid = false;
        if (i < C) {
            valid = seg_valid(cons[i], nmax);
            h = hcnt[i];
            np[i] = valid ? (h <= SEG_SORT_MAX ? h + 1 : 0) : 0;
        }
        bool srt = valid && h >= 2 && h <= SEG_SORT_MAX, wlk = valid && h > SEG_SORT_MAX;
        int a = wave_alloc(srt ? 1 : 0, nsort);
        if (srt) sort_list[a] = (int)i;
        int b = wave_alloc(wlk ? 1 : 0, nwalk);
        if (wlk) walk_list[b] = (int)i;
    }
}

// one workgroup per listed segment: its vertices sorted along it (bitonic
// sort in LDS of (position along the segment, vertex) pairs)
__global__ __launch_bounds__(kBlock) void k_seg_sort(const int2* pts, const int2* cons, const int* list, const int* nlist,
                                                     const int* hcnt, const int* hoff, int* hits)
{
    __shared__ long long s_key[SEG_SORT_MAX];
    __shared__ int s_val[SEG_SORT_MAX];
    int n = *nlist;
    for (int w = blockIdx.x; w < n; w += gridDim.x) {
        int sg = list[w];
        int h = hcnt[sg], o = hoff[sg];
        int2 e = cons[sg];
        int2 A = pts[e.x], B = pts[e.y];
        long long dx = (long long)B.x - A.x, dy = (long long)B.y - A.y;
        int m = 1;
        while (m < h) m <<= 1;
        for (int k = threadIdx.x; k < m; k += blockDim.x) {
            if (k < h) {
                int p = hits[o + k];
                int2 P = pts[p];
                s_key[k] = ((long long)P.x - A.x) * dx + ((long long)P.y - A.y) * dy;
                s_val[k] = p;
            } else {
                s_key[k] = LLONG_MAX;
                s_val[k] = -1;
            }
        }
        __syncthreads();
        for (int size = 2; size <= m; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int k = threadIdx.x; k < m; k += blockDim.x) {
                    int j = k ^ stride;
                    if (j > k) {
                        bool up = (k & size) == 0;
                        long long a = s_key[k], b = s_key[j];
                        if ((a > b) == up) {
                            s_key[k] = b;
                            s_key[j] = a;
                            int t = s_val[k];
                            s_val[k] = s_val[j];
                            s_val[j] = t;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int k = threadIdx.x; k < h; k += blockDim.x) hits[o + k] = s_val[k];
        __syncthreads();
    }
}

__device__ inline int upper_bound_seg(const int* a, int n, long long x)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// one thread per piece of the sorted segments: (previous vertex, next vertex)
// along the segment; an existing edge is marked constrained and stored with
// a negative first vertex
__global__ __launch_bounds__(kBlock) void k_seg_pieces(Cons c, const int2* cons, int C, const int* poff,
                                                       const int* hcnt, const int* hoff, const int* hits,
                                                       long long total, int2* pieces)
{
    const Mesh& m = c.m;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += (long long)gridDim.x * blockDim.x) {
        int sg = upper_bound_seg(poff, C + 1, j) - 1;
        int h = hcnt[sg];
        if (h > SEG_SORT_MAX) continue;  // the walk writes this segment's pieces
        int k = (int)(j - poff[sg]);
        int2 e = cons[sg];
        // the q-th vertex from e.x (hits are sorted from e.x towards e.y)
        auto vtx = [&](int q) { return q == 0 ? e.x : (q == h + 1 ? e.y : hits[hoff[sg] + q - 1]); };
        int u = vtx(k), v = vtx(k + 1);
        int t, s;
        int r = find_wedge(c, u, v, t, s);
        if (r < 0) {
            at_or(c.err, MERR_WALK);
            pieces[j] = make_int2(u, v);
            continue;
        }
        if (r == 1 && get3(m.tri[t], nx3(s)) == v) {
            mark_pair(c, t, pv3(s));
            pieces[j] = make_int2(-(u + 1), v);
        } else {
            pieces[j] = make_int2(u, v);
        }
    }
}

// the walk's pieces for the listed (long) segments: counted into np, then
// written at their offsets
__global__ __launch_bounds__(kBlock) void k_seg_walk(Cons c, const int2* cons, const int* list, const int* nlist,
                                                     int* np, const int* poff, int2* pieces)
{
    int n = *nlist;
    for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < n; w += (long long)gridDim.x * blockDim.x) {
        int sg = list[w];
        int2 e = cons[sg];
        if (poff) split_walk(c, e.x, e.y, pieces + poff[sg]);
        else np[sg] = split_walk(c, e.x, e.y, nullptr);
    }
}

static int split_pieces_cells(Ctx& cx, const Cons& c, const int* cell_end, const int2* cons, int C, int nmax,
                              DBuf<int2>& pieces)
{
    const int2* pts = c.m.pts;
    DBuf<int> off((size_t)C + 1);
    k_seg_ncells<<<grid_for(C), kBlock, 0, cx.st>>>(pts, cons, C, nmax, off.p, c.err);
    GCDT_LAUNCH_CHECK();
    cx.check_err();
    scan_excl_total(cx, off.p, off.p, C);
    long long npairs = cx.get(C_TOTAL);
    DBuf<int> pseg(npairs > 0 ? npairs : 1), pcell(npairs > 0 ? npairs : 1);
    k_seg_cells<<<grid_for(C), kBlock, 0, cx.st>>>(pts, cons, C, nmax, off.p, pseg.p, pcell.p);
    GCDT_LAUNCH_CHECK();
    off.release();
    // vertices strictly inside each segment: count, offsets, positions
    DBuf<int> hcnt((size_t)C + 1), hoff((size_t)C + 1);
    fill_i32(cx, hcnt.p, 0, C);
    k_seg_hits<<<grid_for(npairs), kBlock, 0, cx.st>>>(pts, cell_end, cons, pseg.p, pcell.p, npairs, hcnt.p, nullptr,
                                                       nullptr);
    GCDT_LAUNCH_CHECK();
    scan_excl_total(cx, hcnt.p, hoff.p, C);
    long long nh = cx.get(C_TOTAL);
    DBuf<int> hits(nh > 0 ? nh : 1);
    fill_i32(cx, hcnt.p, 0, C);
    k_seg_hits<<<grid_for(npairs), kBlock, 0, cx.st>>>(pts, cell_end, cons, pseg.p, pcell.p, npairs, hcnt.p, hoff.p,
                                                       hits.p);
    GCDT_LAUNCH_CHECK();
    pseg.release();
    pcell.release();
    // pieces per segment; segments with many vertices keep the walk
    DBuf<int> np((size_t)C + 1), slist(C > 0 ? C : 1), wlist(C > 0 ? C : 1);
    cx.set(C_NQ, 0);
    cx.set(C_NQ2, 0);
    k_seg_npieces<<<grid_for(C), kBlock, 0, cx.st>>>(cons, C, nmax, hcnt.p, np.p, slist.p, cx.dcnt.p + C_NQ, wlist.p,
                                                     cx.dcnt.p + C_NQ2);
    GCDT_LAUNCH_CHECK();
    k_seg_sort<<<std::min(std::max(C, 1), 65536), kBlock, 0, cx.st>>>(pts, cons, slist.p, cx.dcnt.p + C_NQ, hcnt.p,
                                                                      hoff.p, hits.p);
    GCDT_LAUNCH_CHECK();
    cx.read(C_NQ, 2);
    int nwalk = cx.hcnt[C_NQ2];
    if (nwalk > 0) {
        k_seg_walk<<<grid_for(nwalk), kBlock, 0, cx.st>>>(c, cons, wlist.p, cx.dcnt.p + C_NQ2, np.p, nullptr, nullptr);
        GCDT_LAUNCH_CHECK();
        cx.check_err();
    }
    scan_excl_total(cx, np.p, np.p, C);
    int total = cx.get(C_TOTAL);
    pieces.alloc(total > 0 ? total : 1);
    if (total > 0) {
        k_seg_pieces<<<grid_for(total), kBlock, 0, cx.st>>>(c, cons, C, np.p, hcnt.p, hoff.p, hits.p, total, pieces.p);
        GCDT_LAUNCH_CHECK();
    }
    if (nwalk > 0) {
        k_seg_walk<<<grid_for(nwalk), kBlock, 0, cx.st>>>(c, cons, wlist.p, cx.dcnt.p + C_NQ2, nullptr, np.p,
                                                          pieces.p);
        GCDT_LAUNCH_CHECK();
    }
    cx.check_err();
    return total;
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
                              int* touched, const int* cell_end, ConsDefer* defer)
{
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
    c.steiner = defer != nullptr;
    // 1. split into pieces
    DBuf<int2> pieces;
    int np = cell_end ? split_pieces_cells(cx, c, cell_end, cons, C, nmax, pieces)
                      : split_pieces(cx, c, cons, C, nmax, pieces);
    res.pieces = np;
    DBuf<int> dflag;
    if (defer) {
        dflag.alloc(np > 0 ? np : 1);
        fill_i32(cx, dflag.p, 0, np);
        c.defer = dflag.p;
    }
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
        // sum of cav_scratch_ints(K) over the queued pieces (13 K + 20 each)
        long long cap = 13LL * tot + 20LL * nq + 1024;
        if (cap > (1LL << 30)) cap = 1LL << 30;
        scr.ensure((size_t)cap);
        rew.ensure((size_t)tot + 1);
        k_cav_check<<<g, kBlock, 0, cx.st>>>(c, pieces.p, Q, cx.dcnt.p + C_NQ, ncross.p, ep, Q2, cx.dcnt.p + C_NQ2,
                                             wins.p, cx.dcnt.p + C_NWIN, cx.dcnt.p + C_SCR, (long long)scr.n);
        GCDT_LAUNCH_CHECK();
        k_cav_retri<<<g, kBlock, 0, cx.st>>>(c, pieces.p, wins.p, cx.dcnt.p + C_NWIN, scr.p, ep, rew.p,
                                             cx.dcnt.p + C_NREW);
        GCDT_LAUNCH_CHECK();
        // one wave per winner, up to 8192 waves in flight
        const int gw = (int)((std::min<long long>(nq, 8192) * 64 + kBlock - 1) / kBlock);
        k_cav_retri_wave<<<gw, kBlock, 0, cx.st>>>(c, pieces.p, wins.p, cx.dcnt.p + C_NWIN, scr.p, ep, rew.p,
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
    if (defer) {
        defer->pieces = std::move(pieces);
        defer->flag = std::move(dflag);
        defer->np = np;
    }
    return res;
}

}  // namespace gcdt
