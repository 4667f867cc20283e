// Crossing pass of gcdt_solve_steiner on gfx950: the deferred constraint
// pieces are inserted into the CDT of the others by one device thread
// (steiner_bodies.hpp), in piece order, so the crossing vertices are numbered
// the same way on every run with the same deferred set.
#include <algorithm>

#include "mesh.hpp"
#include "steiner_bodies.hpp"

namespace gcdt {

__global__ __launch_bounds__(64) void k_steiner(Stn s, const int2* pieces, const int* list, int n)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int i = 0; i < n; i++) {
        const int2 pc = pieces[list[i]];
        if (!stn_piece(s, pc.x, pc.y)) return;
    }
}

// vertex -> its incident triangle of largest id (vt filled with -1 first):
// the serial pass starts its walks there, and the triangles it splits and
// rewrites follow from that start, so it is the same on every run
__global__ __launch_bounds__(kBlock) void k_stn_vt(const int3* tri, int T, int* vt)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        const int3 t = tri[i];
        atomicMax(&vt[t.x], (int)i);
        atomicMax(&vt[t.y], (int)i);
        atomicMax(&vt[t.z], (int)i);
    }
}

// crossing points -> gcdt_hpoint, their pieces in caller ids
__global__ __launch_bounds__(kBlock) void k_steiner_out(const HPt* hp, const int2* sup, int S, const int* perm, int n,
                                                        gcdt_hpoint* out)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < S; i += (long long)gridDim.x * blockDim.x) {
        const HPt h = hp[i];
        gcdt_hpoint o;
        o.x_lo = h.X.w[0];
        o.x_hi = (int64_t)h.X.w[1];
        o.y_lo = h.Y.w[0];
        o.y_hi = (int64_t)h.Y.w[1];
        o.w = h.W;
        const int2 a = sup[STN_SUPS * i], b = sup[STN_SUPS * i + 1];
        const int v[4] = {a.x, a.y, b.x, b.y};
        for (int j = 0; j < 4; j++) o.seg[j] = v[j] < n ? perm[v[j]] : v[j];
        out[i] = o;
    }
}

int steiner_pass(Ctx& c, const Mesh& m, int T, int tcap, int nv0, const ConsDefer& d, int scap, gcdt_hpoint* out,
                 const int* perm, int n, int* S)
{
    *S = 0;
    if (d.np <= 0 || !d.flag.p) return T;
    DBuf<int> list(d.np);
    const int nd = flags_to_list(c, d.flag.p, d.np, list.p);
    if (nd == 0) return T;
    if (scap <= 0) throw Error(GCDT_ERR_CAPACITY, "constraints cross and max_steiner is 0");
    const long long nv = (long long)nv0 + scap;
    DBuf<int> vt((size_t)nv);
    fill_i32(c, vt.p, -1, nv);
    k_stn_vt<<<grid_for(T), kBlock, 0, c.st>>>(m.tri, T, vt.p);
    GCDT_LAUNCH_CHECK();
    DBuf<HPt> hp((size_t)scap);
    DBuf<int2> sup((size_t)scap * STN_SUPS);
    // queue and stacks of the serial pass (vertex pairs)
    const long long scr_n = std::min<long long>(std::max<long long>(1LL << 20, 4LL * T), 1LL << 26);
    DBuf<int> scr((size_t)scr_n);
    c.set(C_STN_T, T);
    c.set(C_STN_S, 0);
    Stn s;
    s.tri = m.tri;
    s.adj = m.adj;
    s.cflag = m.cflag;
    s.pts = m.pts;
    s.nv0 = nv0;
    s.hp = hp.p;
    s.sup = sup.p;
    s.vt = vt.p;
    s.scr = scr.p;
    s.scr_cap = scr_n;
    s.ctr = c.dcnt.p + C_STN_T;
    s.tcap = tcap;
    s.scap = scap;
    s.err = c.dcnt.p + C_ERR;
    {
        ProfScope ps("steiner", c.st, nd);
        k_steiner<<<1, 64, 0, c.st>>>(s, d.pieces.p, list.p, nd);
        GCDT_LAUNCH_CHECK();
    }
    c.read(0, C_NUM);
    const int cap_hit = c.hcnt[C_STN_S] >= scap && (c.hcnt[C_ERR] & MERR_CAPACITY);
    if (cap_hit) throw Error(GCDT_ERR_CAPACITY, "more than max_steiner crossing points");
    if (c.hcnt[C_ERR]) c.check_err();
    *S = c.hcnt[C_STN_S];
    if (*S > 0) {
        k_steiner_out<<<grid_for(*S), kBlock, 0, c.st>>>(hp.p, sup.p, *S, perm, n, out);
        GCDT_LAUNCH_CHECK();
    }
    return c.hcnt[C_STN_T];
}

}  // namespace gcdt
