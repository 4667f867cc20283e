// Host execution of the solve's per-item kernel bodies (mesh_bodies.hpp,
// cons_bodies.hpp), one item at a time in seeded shuffled orders: a CPU test
// of the claim, flip and fix protocols (test infrastructure, not product
// code).  The GPU's list mechanics (staging, claimed-triangle lists, tails)
// are not modelled; the bodies' results must still form a valid CDT.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <vector>

#include <sys/mman.h>

namespace gcdt {
__host__ inline unsigned long long ld_u64(const unsigned long long* p) { return *p; }
__host__ inline unsigned long long at_max(unsigned long long* p, unsigned long long v) { auto o = *p; if (v > o) *p = v; return o; }
__host__ inline unsigned long long at_min(unsigned long long* p, unsigned long long v) { auto o = *p; if (v < o) *p = v; return o; }
__host__ inline int at_max(int* p, int v) { int o = *p; if (v > o) *p = v; return o; }
__host__ inline int at_min(int* p, int v) { int o = *p; if (v < o) *p = v; return o; }
__host__ inline int at_or(int* p, int v) { int o = *p; *p |= v; return o; }
__host__ inline int at_add(int* p, int v) { int o = *p; *p += v; return o; }
__host__ inline unsigned long long at_max_wg(unsigned long long* p, unsigned long long v) { return at_max(p, v); }
__host__ inline unsigned long long ld_u64_wg(const unsigned long long* p) { return *p; }
}  // namespace gcdt

#include "../gcdt_amd/csrc/cons_bodies.hpp"

using namespace gcdt;

static const int GR = (1 << 30) - 1;
static int g_err = 0;
static unsigned g_epoch = 0;
static unsigned g_seed = 0;  // item-order shuffles
static int g_fused = 1;      // packed rounds settle flips inside the next test
static int g_two = 0;        // 1: both triangles of an edge test it in every round (else only short lists)
// 1: a proposer whose record another proposal's flip has rewritten earlier in
// the same apply pass reads it torn -- vertices, neighbours and coordinates
// new, meta/stamp/claim old -- as a GPU load of the record can when it runs
// concurrently with that flip's four 16-byte stores
static int g_torn = 0;

static void shuffle(std::vector<int>& v, unsigned salt)
{
    for (size_t i = v.size(); i > 1; i--) {
        size_t j = (size_t)(hash32((uint32_t)(i * 2654435761u) ^ hash32(salt ^ g_seed)) % i);
        std::swap(v[i - 1], v[j]);
    }
}

struct DAG {
    std::vector<int3> tri, adj;
    std::vector<int> sons;
    std::vector<unsigned long long> claim;
};

static int insert_rounds(std::vector<int2>& pts, DAG& d, std::vector<int>& loc, std::vector<int> pend, int ntri, int* rounds)
{
    int cap = (int)d.tri.size();
    Ins s;
    s.m = Mesh{pts.data(), d.tri.data(), d.adj.data(), d.sons.data(), nullptr};
    s.loc = loc.data();
    s.claim = d.claim.data();
    static std::vector<unsigned long long> pk;
    pk.assign(pts.size(), 0);
    s.pkey = pk.data();
    s.cap = cap;
    s.err = &g_err;
    s.pivot = 1;
    int r = 0;
    while (!pend.empty()) {
        r++;
        unsigned ep = ++g_epoch;
        for (int p : pend) ins_claim(s, p, ep);
        int tbase = ntri;
        // out-of-order slot allocation, as blocks append in any order
        std::vector<int> order(pend);
        shuffle(order, ep);
        for (int p : order) {
            int need = ins_need(s, p, ep);
            if (!need) continue;
            if (ntri + need > cap) { g_err |= 1; return -1; }
            ins_split(s, p, ntri);
            ntri += need;
        }
        for (int c = tbase; c < ntri; c++) ins_fix(s, c);
        std::vector<int> nx;
        for (int p : pend)
            if (ins_relocate(s, p)) nx.push_back(p);
        pend.swap(nx);
        if (g_err) return -1;
    }
    *rounds = r;
    return ntri;
}

static long long flip_rounds(const Mesh& m, int T, int* rounds)
{
    std::vector<unsigned long long> claim(T, 0);
    std::vector<int> stamp(T, 0), partner(T, 0), cand(T, -1), qround(T, 0), mark(T, 0);
    Flip f;
    f.m = m;
    f.claim = claim.data();
    f.stamp = stamp.data();
    f.partner = partner.data();
    f.cand = cand.data();
    f.mark = mark.data();
    f.err = &g_err;
    std::vector<int> act(T);
    for (int i = 0; i < T; i++) act[i] = i;
    long long flips = 0;
    int r = 0;
    while (!act.empty()) {
        r++;
        unsigned ep = ++g_epoch;
        for (int t : act) flip_test(f, t, ep, r == 1, !g_two && act.size() > 64);
        std::vector<int> nx, fl;
        for (int t : act) {
            int n = -1;
            int res = flip_apply(f, t, ep, &n);
            if (res == 1 || res == 2) nx.push_back(t);
            if (res == 2) nx.push_back(n);
            if (res == 2) { fl.push_back(t); fl.push_back(n); }
        }
        for (int t : nx) {
            if (qround[t] == (int)ep) return -1;  // listed twice
            qround[t] = (int)ep;
            mark[t] = (int)ep + 1;
        }
        for (int t : fl) flip_fix(f, t, ep);
        flips += fl.size() / 2;
        act.swap(nx);
        if (r > 10000000) return -1;
    }
    *rounds = r;
    return flips;
}

// Packed-record rounds over the whole mesh, as the GPU runs them: the test of
// round r settles round r-1's flips (flip_fixtest_p) unless g_fused is 0, in
// which case flip_fix_p runs after every apply.  Tests and applies visit the
// list in shuffled orders, so retries meet flipped neighbours both before and
// after those neighbours' own settling.
// Records in a mapping whose first page is inaccessible and holds no record,
// so a read of rec[-1] (a neighbour id of -1 used as an index) faults.
struct GuardedRecs {
    char* base = nullptr;
    size_t bytes = 0;
    FRec* r = nullptr;
    explicit GuardedRecs(int T)
    {
        const size_t page = 4096;
        bytes = page + ((size_t)std::max(T, 1) * sizeof(FRec) + page - 1) / page * page;
        base = (char*)mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        if (base == MAP_FAILED) abort();
        mprotect(base, page, PROT_NONE);
        r = (FRec*)(base + page);
    }
    ~GuardedRecs() { munmap(base, bytes); }
    FRec& operator[](long long i) { return r[i]; }
};

static long long flip_rounds_p(const Mesh& m, int T, int* rounds)
{
    GuardedRecs rec(T);
    for (int t = 0; t < T; t++) {
        int3 v = m.tri[t];
        rec_store(&rec[t], v, m.adj[t], m.pts[v.x], m.pts[v.y], m.pts[v.z],
                  (m.cflag ? m.cflag[t] & 7 : 0) | FREC_NOCAND | FREC_NOPARTNER, 0, 0ull);
    }
    std::vector<int> qround(T, 0);
    FlipP f{rec.r, &g_err};
    std::vector<int> act(T);
    for (int i = 0; i < T; i++) act[i] = i;
    long long flips = 0;
    int r = 0;
    while (!act.empty()) {
        r++;
        unsigned ep = ++g_epoch;
        std::vector<int> ord(act);
        shuffle(ord, ep);
        for (int t : ord) {
            if (g_fused) flip_fixtest_p(f, t, ep, r == 1, !g_two && act.size() > 64);
            else flip_test_p(f, t, ep, r == 1, !g_two && act.size() > 64);
        }
        std::vector<int> nx, fl;
        shuffle(ord, ~ep);
        std::vector<FRec> before;
        if (g_torn) before.assign(rec.r, rec.r + T);
        for (int t : ord) {
            int n = -1;
            int res;
            if (g_torn && memcmp(&rec[t], &before[t], sizeof(FRec)) != 0) {
                // rewritten by another proposal's flip in this pass
                const FRec now = rec[t];
                memcpy((char*)&rec[t] + 48, (const char*)&before[t] + 48, 16);
                res = flip_apply_p(f, t, ep, &n);
                if (res == 2) return -1;  // t flipped from a torn read
                rec[t] = now;
            } else {
                res = flip_apply_p(f, t, ep, &n);
            }
            if (res == 1) rec[t].meta = rec_with_mark(rec[t].meta, (int)ep + 1);
            if (res == 1 || res == 2) nx.push_back(t);
            if (res == 2) {
                nx.push_back(n);
                fl.push_back(t);
                fl.push_back(n);
            }
        }
        for (int t : nx) {
            if (qround[t] == (int)ep || !rec_marked(rec[t].meta, ep + 1)) return -1;  // listed twice / unmarked
            qround[t] = (int)ep;
        }
        if (!nx.empty() && fl.empty()) return -1;  // a round without progress
        // fused rounds: the last round's flips have no later test to settle them
        if (!g_fused || nx.empty())
            for (int t : fl) flip_fix_p(f, t, ep);
        flips += fl.size() / 2;
        act.swap(nx);
        if (r > 10000000) return -1;
    }
    for (int t = 0; t < T; t++) {
        m.tri[t] = rec[t].v;
        m.adj[t] = rec[t].a;
        if (m.cflag) m.cflag[t] = rec[t].meta & 7;
        int2 a = m.pts[rec[t].v.x], b = m.pts[rec[t].v.y], c = m.pts[rec[t].v.z];
        if (a.x != rec[t].p[0].x || a.y != rec[t].p[0].y || b.x != rec[t].p[1].x || b.y != rec[t].p[1].y ||
            c.x != rec[t].p[2].x || c.y != rec[t].p[2].y)
            return -1;  // a record's coordinates disagree with its vertices
    }
    *rounds = r;
    return flips;
}

static int filter(const DAG& d, int T, std::vector<int3>& tri2, std::vector<int3>& adj2)
{
    std::vector<int> off(T + 1, 0);
    for (int i = 0; i < T; i++) off[i + 1] = off[i] + (d.sons[i] == -1);
    int n = off[T];
    tri2.resize(n);
    adj2.resize(n);
    for (int i = 0; i < T; i++) {
        if (d.sons[i] != -1) continue;
        tri2[off[i]] = d.tri[i];
        int3 a = d.adj[i];
        a.x = a.x >= 0 ? off[a.x] : -1;
        a.y = a.y >= 0 ? off[a.y] : -1;
        a.z = a.z >= 0 ? off[a.z] : -1;
        adj2[off[i]] = a;
    }
    return n;
}

static void dag_init(DAG& d, int cap, int N)
{
    d.tri.assign(cap, make_int3(-9, -9, -9));
    d.adj.assign(cap, make_int3(-1, -1, -1));
    d.sons.assign(cap, -1);
    d.claim.assign(cap, 0);
    d.tri[0] = make_int3(N, N + 1, N + 3);
    d.tri[1] = make_int3(N + 1, N + 2, N + 3);
    d.adj[0] = make_int3(1, -1, -1);
    d.adj[1] = make_int3(-1, 0, -1);
}

// Full solve on the host: sample mesh, located insertion, packed flip rounds,
// constraint cavities and the constrained flip pass.  Returns 0 or an error
// code (the hundreds say which stage failed); tris_out holds 2N+2 triangles.
extern "C" void host_set_torn(int on) { g_torn = on; }

extern "C" int host_solve(const int* xy, int N, const int* consp, int C, int* tris_out, long long* stats,
                          unsigned seed, int fused, int two_testers)
{
    g_err = 0;
    g_seed = seed;
    g_fused = fused;
    g_two = two_testers;
    std::vector<int2> pts(N + 4);
    for (int i = 0; i < N; i++) pts[i] = make_int2(xy[2 * i], xy[2 * i + 1]);
    pts[N] = make_int2(0, 0);
    pts[N + 1] = make_int2(GR, 0);
    pts[N + 2] = make_int2(GR, GR);
    pts[N + 3] = make_int2(0, GR);
    // sample: smallest id per 256^2 cell
    std::vector<int> cellmin(1 << 16, INT_MAX);
    for (int i = 0; i < N; i++) {
        int c = ((pts[i].x >> 22) << 8) + (pts[i].y >> 22);
        cellmin[c] = std::min(cellmin[c], i);
    }
    std::vector<int> sid;
    for (int c = 0; c < (1 << 16); c++)
        if (cellmin[c] != INT_MAX) sid.push_back(cellmin[c]);
    int S = (int)sid.size();
    std::vector<int> loc(N + 4, -1);
    for (int p : sid) {
        int o = orient2d(make_int2(GR, 0), make_int2(0, GR), pts[p]);
        loc[p] = o > 0 ? 3 : (o < 0 ? 7 : 0);
    }
    DAG A;
    dag_init(A, 4 * S + 8, N);
    int ra = 0;
    int ntA = insert_rounds(pts, A, loc, sid, 2, &ra);
    if (ntA < 0) return 100 + g_err;
    std::vector<int3> tS, aS;
    int Ts = filter(A, ntA, tS, aS);
    int fr = 0;
    long long fl1 = flip_rounds_p(Mesh{pts.data(), tS.data(), aS.data(), nullptr, nullptr}, Ts, &fr);
    if (fl1 < 0 || g_err) return 200 + g_err;
    // locate the rest by brute force over the sample triangles
    std::vector<char> isv(N, 0);
    for (auto& t : tS) {
        if (t.x < N) isv[t.x] = 1;
        if (t.y < N) isv[t.y] = 1;
        if (t.z < N) isv[t.z] = 1;
    }
    DAG B;
    B.tri.assign(4 * (size_t)N + 8, make_int3(-9, -9, -9));
    B.adj.assign(4 * (size_t)N + 8, make_int3(-1, -1, -1));
    B.sons.assign(4 * (size_t)N + 8, -1);
    B.claim.assign(4 * (size_t)N + 8, 0);
    for (int i = 0; i < Ts; i++) { B.tri[i] = tS[i]; B.adj[i] = aS[i]; }
    std::vector<int> pend;
    for (int p = 0; p < N; p++) {
        if (isv[p]) { loc[p] = -1; continue; }
        int found = -1;
        for (int t = 0; t < Ts; t++) {
            int r = classify(pts[tS[t].x], pts[tS[t].y], pts[tS[t].z], pts[p]);
            if (r >= 0) { found = t * 4 + r; break; }
            if (r == -2) return 300;
        }
        if (found < 0) return 301;
        loc[p] = found;
        pend.push_back(p);
    }
    int rb = 0;
    int ntB = insert_rounds(pts, B, loc, pend, Ts, &rb);
    if (ntB < 0) return 400 + g_err;
    std::vector<int3> tri, adj;
    int T = filter(B, ntB, tri, adj);
    if (T != 2 * N + 2) return 500;
    int f2 = 0;
    long long fl2 = flip_rounds_p(Mesh{pts.data(), tri.data(), adj.data(), nullptr, nullptr}, T, &f2);
    if (fl2 < 0 || g_err) return 600 + g_err;
    long long pieces_n = 0, crounds = 0;
    if (C > 0) {
        std::vector<int> cflag(T, 0), vt(N + 4, -1), stamp(T, 0), cavof(T, 0);
        std::vector<unsigned long long> owner(T, ~0ull);
        for (int t = 0; t < T; t++) { vt[tri[t].x] = t; vt[tri[t].y] = t; vt[tri[t].z] = t; }
        Cons c;
        c.m = Mesh{pts.data(), tri.data(), adj.data(), nullptr, cflag.data()};
        c.vt = vt.data();
        c.owner = owner.data();
        c.stamp = stamp.data();
        c.cavof = cavof.data();
        c.err = &g_err;
        std::vector<int2> cons(C);
        for (int i = 0; i < C; i++) cons[i] = make_int2(consp[2 * i], consp[2 * i + 1]);
        std::vector<int> off(C + 1, 0);
        for (int i = 0; i < C; i++) off[i + 1] = off[i] + split_walk(c, cons[i].x, cons[i].y, nullptr);
        if (g_err) return 700 + g_err;
        std::vector<int2> pieces(off[C]);
        for (int i = 0; i < C; i++) split_walk(c, cons[i].x, cons[i].y, pieces.data() + off[i]);
        pieces_n = off[C];
        std::vector<int> Q;
        for (int i = 0; i < off[C]; i++)
            if (pieces[i].x >= 0) Q.push_back(i);
        while (!Q.empty()) {
            crounds++;
            unsigned ep = ++g_epoch;
            std::vector<int> ncross(Q.size());
            long long tot = 0;
            for (size_t i = 0; i < Q.size(); i++) {
                int pi = Q[i];
                unsigned long long key = cav_key(ep, pi);
                int k = cav_walk(c, pieces[pi].x, pieces[pi].y, true, nullptr, [&](int t) { at_min(&owner[t], key); });
                if (k < 0) return 800 + g_err;
                ncross[i] = k;
                tot += k;
            }
            std::vector<int> scr(13 * tot + 20 * Q.size() + 64);
            std::vector<Win> wins;
            std::vector<int> Q2;
            int cursor = 0;
            for (size_t i = 0; i < Q.size(); i++) {
                int pi = Q[i];
                if (ncross[i] == 0) continue;
                unsigned long long key = cav_key(ep, pi);
                bool all = true;
                int t0 = -1;
                cav_walk(c, pieces[pi].x, pieces[pi].y, false, &t0, [&](int t) { all &= owner[t] == key; });
                if (all) {
                    wins.push_back(Win{pi, t0, ncross[i], cursor});
                    cursor += (int)cav_scratch_ints(ncross[i]);
                } else {
                    Q2.push_back(pi);
                }
            }
            std::vector<int> rew;
            for (size_t wi = 0; wi < wins.size(); wi++) {
                if (!cav_retri_one(c, pieces.data(), wins.data(), (int)wi, scr.data(), ep)) return 900 + g_err;
                for (int j = 0; j < wins[wi].K; j++) rew.push_back(scr[wins[wi].off + j]);
            }
            for (int t : rew) cav_fix_one(c, wins.data(), scr.data(), t, ep);
            if (g_err) return 1000 + g_err;
            if (wins.empty() && !Q2.empty()) return 1100;
            Q.swap(Q2);
        }
        int f3 = 0;
        long long fl3 = flip_rounds(c.m, T, &f3);
        if (fl3 < 0 || g_err) return 1200 + g_err;
        fl2 += fl3;
        f2 += f3;
    }
    for (int t = 0; t < T; t++) {
        tris_out[3 * t] = tri[t].x;
        tris_out[3 * t + 1] = tri[t].y;
        tris_out[3 * t + 2] = tri[t].z;
    }
    if (stats) {
        stats[0] = T;
        stats[1] = ra + rb;
        stats[2] = fr + f2;
        stats[3] = fl1 + fl2;
        stats[4] = pieces_n;
        stats[5] = crounds;
        stats[6] = S;
    }
    return 0;
}
