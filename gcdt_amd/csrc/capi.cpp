// C ABI of the library: runtime helpers, the reference's entry points and the
// one-call solver.  Every function catches exceptions at the boundary and
// returns a GCDT_* status.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mesh.hpp"

using namespace gcdt;

static thread_local std::string g_last_error;

namespace gcdt {
// records the calling thread's error message and returns the status
int fail(int status, const std::string& msg)
{
    g_last_error = msg;
    return status;
}
}  // namespace gcdt

#define API_BEGIN try {
#define API_END                                             \
    }                                                       \
    catch (const Error& e) { return fail(e.status, e.what()); } \
    catch (const std::exception& e) { return fail(GCDT_ERR_INTERNAL, e.what()); }

// hipEvent-timed region on a stream
struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t st;
    // inside an arena call no events are created: the host clock from here
    // to the stream's drain in ms()
    std::chrono::steady_clock::time_point h0;
    explicit Timer(hipStream_t s) : st(s)
    {
        if (g_arena) {
            h0 = std::chrono::steady_clock::now();
            return;
        }
        GCDT_HIP(hipEventCreate(&a));
        GCDT_HIP(hipEventCreate(&b));
        GCDT_HIP(hipEventRecord(a, st));
    }
    double ms()
    {
        if (!a) {
            GCDT_HIP(hipStreamSynchronize(st));
            return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
        }
        GCDT_HIP(hipEventRecord(b, st));
        GCDT_HIP(hipEventSynchronize(b));
        float f = 0;
        GCDT_HIP(hipEventElapsedTime(&f, a, b));
        return f;
    }
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

extern "C" {

const char* gcdt_last_error(void) { return g_last_error.c_str(); }

int gcdt_device_count(int* count)
{
    API_BEGIN
    int n = 0;
    GCDT_HIP(hipGetDeviceCount(&n));
    if (count) *count = n;
    return GCDT_OK;
    API_END
}

int gcdt_device_name(char* buf, int len)
{
    API_BEGIN
    if (!buf || len <= 0) return fail(GCDT_ERR_INVALID, "bad buffer");
    int dev = 0;
    GCDT_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    GCDT_HIP(hipGetDeviceProperties(&prop, dev));
    snprintf(buf, (size_t)len, "%s:%s:cu%d", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return GCDT_OK;
    API_END
}

int gcdt_synchronize(void)
{
    API_BEGIN
    GCDT_HIP(hipDeviceSynchronize());
    return GCDT_OK;
    API_END
}

int gcdt_alloc(void** ptr, size_t bytes)
{
    API_BEGIN
    void* p = getMemCacheRef().getMemByAbtrtSize(bytes);
    if (!p) return fail(GCDT_ERR_OOM, "pool allocation failed");
    *ptr = p;
    return GCDT_OK;
    API_END
}

int gcdt_release(void* ptr, size_t bytes)
{
    API_BEGIN
    getMemCacheRef().releaseAbtrtMem(ptr, bytes);
    return GCDT_OK;
    API_END
}

int gcdt_pool_init(void)
{
    API_BEGIN
    getMemCacheRef().initAbtrtMem();
    noalloc_setup();
    return GCDT_OK;
    API_END
}

int gcdt_pool_stats(size_t* in_use, size_t* cached, size_t* peak)
{
    MemCache& m = getMemCacheRef();
    if (in_use) *in_use = m.bytesInUse();
    if (cached) *cached = m.bytesCached();
    if (peak) *peak = m.peakBytes();
    return GCDT_OK;
}

int gcdt_pool_set_budget(size_t bytes)
{
    getMemCacheRef().setBudget(bytes);
    return GCDT_OK;
}

int gcdt_pool_reset_peak(void)
{
    getMemCacheRef().resetPeak();
    return GCDT_OK;
}

int gcdt_pool_trim(void)
{
    getMemCacheRef().releaseCache();
    return GCDT_OK;
}

int gcdt_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    API_BEGIN
    GCDT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return GCDT_OK;
    API_END
}

int gcdt_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    API_BEGIN
    GCDT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GCDT_OK;
    API_END
}

int gcdt_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
    API_BEGIN
    GCDT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return GCDT_OK;
    API_END
}

int gcdt_memset(void* dst, int value, size_t bytes)
{
    API_BEGIN
    GCDT_HIP(hipMemset(dst, value, bytes));
    return GCDT_OK;
    API_END
}

size_t gcdt_scratch_bytes(int64_t n) { return sizeof(int) * (size_t)scan_scratch_ints(n); }

int gcdt_stream_create(void** stream)
{
    API_BEGIN
    if (!stream) return fail(GCDT_ERR_INVALID, "null stream pointer");
    hipStream_t s = nullptr;
    GCDT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return GCDT_OK;
    API_END
}

int gcdt_stream_destroy(void* stream)
{
    API_BEGIN
    if (stream) GCDT_HIP(hipStreamDestroy((hipStream_t)stream));
    return GCDT_OK;
    API_END
}

int gcdt_stream_synchronize(void* stream)
{
    API_BEGIN
    GCDT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return GCDT_OK;
    API_END
}

int gcdt_memcpy_h2d_async(void* dst, const void* src, size_t bytes, void* stream)
{
    API_BEGIN
    GCDT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return GCDT_OK;
    API_END
}

int gcdt_memcpy_d2h_async(void* dst, const void* src, size_t bytes, void* stream)
{
    API_BEGIN
    GCDT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    return GCDT_OK;
    API_END
}

int gcdt_prof_enable(int on)
{
    prof_enable(on != 0);
    return GCDT_OK;
}

int gcdt_prof_only(const char* kernel)
{
    prof_only(kernel);
    return GCDT_OK;
}

int gcdt_prof_reset(void)
{
    prof_reset();
    return GCDT_OK;
}

int gcdt_prof_read(const char* kernel, double* total_ms, int64_t* launches, int64_t* items)
{
    long long l = 0, it = 0;
    double ms = 0;
    if (!prof_read(kernel, &ms, &l, &it)) return fail(GCDT_ERR_INVALID, "kernel not profiled");
    if (total_ms) *total_ms = ms;
    if (launches) *launches = l;
    if (items) *items = it;
    return GCDT_OK;
}

int gcdt_prof_names(char* buf, int cap) { return prof_names(buf, cap); }

}  // extern "C"

// ------------------------------------------------------------------ the solve

namespace gcdt {

// Full pipeline on device arrays; pts holds N+4 entries (corners written here).
void solve_impl(Ctx& c, const int2* pts_in, int N, const int2* cons_in, int C, int3* tris_out, int3* adj_out,
                SolveOut& out, int* cflag_out, const SteinerArgs* steiner)
{
    gcdt_stats& st = out.st;
    Timer total(c.st);
    check_points(c, pts_in, N);
    // work in Morton order for locality; ids are mapped back at the end
    DBuf<int2> pts_buf((size_t)N + 4);
    DBuf<int> perm(N > 0 ? N : 1), inv(N > 0 ? N : 1);
    DBuf<int2> cons_buf(C > 0 ? C : 1);
    int2* pts = pts_buf.p;
    DBuf<int> cell_end((size_t)(1 << 20) + 1);
    morton_order(c, pts_in, N, pts, perm.p, inv.p, cell_end.p);
    relabel_pairs(c, cons_in, C, inv.p, N, cons_buf.p);
    const int2* cons = cons_buf.p;
    write_corners(c, pts, N);
    // ---- 1. sample + sample mesh
    Timer t1(c.st);
    DBuf<int> sid(1 << 16);
    int S = sample_points_morton(c, pts, cell_end.p, sid.p);
    // cell_end stays for the constraints' split pass
    DBuf<int> loc((size_t)N + 4), pend((size_t)(N > 1 ? N : 1)), pend2((size_t)(N > 1 ? N : 1));
    DBuf<unsigned long long> pkey((size_t)N + 4);
    long long capA = 4LL * S + 8;
    DBuf<int3> tA(capA), aA(capA);
    DBuf<int> sA(capA);
    DBuf<unsigned long long> clA(capA);
    init_square(c, nullptr, tA.p, aA.p, N);
    fill_i32(c, sA.p, -1, capA);
    fill_u64(c, clA.p, 0, capA);
    locate_two_ids(c, pts, sid.p, S, loc.p, pend.p);
    Ins ia;
    ia.m = Mesh{pts, tA.p, aA.p, sA.p, nullptr};
    ia.loc = loc.p;
    ia.claim = clA.p;
    ia.pkey = pkey.p;
    ia.cap = (int)capA;
    ia.err = c.dcnt.p + C_ERR;
    int rs = 0;
    int ntA = insert_rounds(c, ia, pend, pend2, S, 2, &rs);
    long long capB = 4LL * N + 8;
    DBuf<int3> tB(capB), aB(capB);
    int Ts = filter_leaves_morton(c, pts, tA.p, aA.p, sA.p, ntA, tB.p, aB.p, capB);
    tA.release();
    aA.release();
    sA.release();
    clA.release();
    validate_mesh(c, "the sample mesh's insertion", Mesh{pts, tB.p, aB.p, nullptr, nullptr}, N + 4, Ts, false);
    int fr = 0;
    st.flips += flip_rounds(c, Mesh{pts, tB.p, aB.p, nullptr, nullptr}, Ts, nullptr, 0, &fr);
    validate_mesh(c, "the sample mesh's flips", Mesh{pts, tB.p, aB.p, nullptr, nullptr}, N + 4, Ts, true);
    st.ms_sample = t1.ms();
    // ---- 2. bitmap over the sample Delaunay mesh
    Timer t2(c.st);
    int *bo = nullptr, *bt = nullptr, nbt = 0;
    build_bitmap(c, pts, tB.p, Ts, &bo, &bt, &nbt);
    st.ms_bitmap = t2.ms();
    // ---- 3. locate + insert the rest
    Timer t3(c.st);
    DBuf<int> isv(N > 0 ? N : 1);
    fill_i32(c, isv.p, 0, N);
    int np = locate_bitmap(c, pts, N, tB.p, Ts, bo, bt, loc.p, isv.p, pend.p);
    dev_give(bo, sizeof(int) * ((1 << 20) + 1));
    dev_give(bt, sizeof(int) * (size_t)(nbt > 0 ? nbt : 1));
    isv.release();
    DBuf<int> sB(capB);
    DBuf<unsigned long long> clB(capB);
    fill_i32(c, sB.p, -1, capB);
    fill_u64(c, clB.p, 0, capB);
    Ins ib;
    ib.m = Mesh{pts, tB.p, aB.p, sB.p, nullptr};
    ib.loc = loc.p;
    ib.claim = clB.p;
    ib.pkey = pkey.p;
    ib.cap = (int)capB;
    ib.err = c.dcnt.p + C_ERR;
    int ri = 0;
    int ntB = insert_rounds(c, ib, pend, pend2, np, Ts, &ri);
    st.insert_rounds = rs + ri;
    clB.release();
    pkey.release();
    loc.release();
    pend.release();
    pend2.release();
    // leaves in Morton order, written straight into the packed flip records
    DBuf<FRec> rec(2LL * N + 2);
    int T = filter_leaves_morton_rec(c, pts, tB.p, aB.p, sB.p, ntB, rec.p, 2LL * N + 2);
    tB.release();
    aB.release();
    sB.release();
    if (T != 2 * N + 2) throw Error(GCDT_ERR_INTERNAL, "leaf count " + std::to_string(T) + " != 2N+2");
    st.ms_insert = t3.ms();
    // ---- 4. Delaunay flips (unpacked into tris_out / adj_out)
    Timer t4(c.st);
    int f2 = 0;
    st.flips += flip_rounds_records(c, rec.p, Mesh{pts, tris_out, adj_out, nullptr, nullptr}, T, &f2);
    validate_mesh(c, "the flips", Mesh{pts, tris_out, adj_out, nullptr, nullptr}, N + 4, T, true);
    rec.release();
    st.flip_rounds = fr + f2;
    st.ms_flip = t4.ms();
    // ---- 5. constraints
    Timer t5(c.st);
    if (cflag_out) fill_i32(c, cflag_out, 0, T);
    if (C > 0) {
        DBuf<int> cflag_buf, touched(T);
        int* cflag = cflag_out;
        // the crossing pass adds two triangles per crossing vertex
        const long long tcap = steiner ? 2LL * ((long long)N + steiner->scap) + 2 : T;
        if (!cflag) {
            cflag_buf.alloc((size_t)tcap);
            cflag = cflag_buf.p;
            fill_i32(c, cflag, 0, T);
        }
        fill_i32(c, touched.p, 0, T);
        Mesh m{pts, tris_out, adj_out, nullptr, cflag};
        ConsDefer defer;
        ConsResult cr = insert_constraints(c, m, T, N + 4, N, cons, C, touched.p, cell_end.p,
                                           steiner ? &defer : nullptr);
        st.constraint_rounds = cr.rounds;
        st.segments = cr.pieces;
        st.splits = cr.pieces - C;
        // only edges of rewritten triangles can be non-Delaunay now
        DBuf<int> tl(T);
        int nt = flags_to_list(c, touched.p, T, tl.p);
        if (nt > 0) {
            int f3 = 0;
            st.flips += flip_rounds(c, m, T, tl.p, nt, &f3);
            st.flip_rounds += f3;
        }
        validate_mesh(c, "the constraints", m, N + 4, T, true);
        if (steiner) {
            tl.release();
            touched.release();
            T = steiner_pass(c, m, T, (int)tcap, N + 4, defer, steiner->scap, steiner->out, perm.p, N, &out.S);
        }
    }
    unmap_tris(c, tris_out, T, perm.p, N);
    st.ms_constraints = t5.ms();
    st.ms_total = total.ms();
    st.num_triangles = T;
    out.T = T;
}

}  // namespace gcdt

extern "C" int gcdt_solve(const gcdt_int2* points, int N, const gcdt_int2* cons, int C, gcdt_int3* tris_out,
                          gcdt_int3* adj_out, int* num_tris, gcdt_stats* stats, void* stream)
{
    API_BEGIN
    if (N < 0 || C < 0 || (N > 0 && !points) || (C > 0 && !cons) || !tris_out)
        return fail(GCDT_ERR_INVALID, "bad arguments");
    if (N > (1 << 28)) return fail(GCDT_ERR_INVALID, "too many points");
    hipStream_t st = (hipStream_t)stream;
    Ctx c(st);
    DBuf<int3> adj_tmp;
    int3* adj = (int3*)adj_out;
    if (!adj) {
        adj_tmp.alloc(2LL * N + 2);
        adj = adj_tmp.p;
    }
    SolveOut out;
    solve_impl(c, (const int2*)points, N, (const int2*)cons, C, (int3*)tris_out, adj, out);
    if (num_tris) *num_tris = out.T;
    if (stats) *stats = out.st;
    return GCDT_OK;
    API_END
}

extern "C" int gcdt_solve_flags(const gcdt_int2* points, int N, const gcdt_int2* cons, int C, gcdt_int3* tris_out,
                                gcdt_int3* adj_out, int* cflag_out, int* num_tris, gcdt_stats* stats, void* stream)
{
    API_BEGIN
    if (N < 0 || C < 0 || (N > 0 && !points) || (C > 0 && !cons) || !tris_out || !adj_out || !cflag_out)
        return fail(GCDT_ERR_INVALID, "bad arguments");
    if (N > (1 << 28)) return fail(GCDT_ERR_INVALID, "too many points");
    Ctx c((hipStream_t)stream);
    SolveOut out;
    solve_impl(c, (const int2*)points, N, (const int2*)cons, C, (int3*)tris_out, (int3*)adj_out, out, cflag_out);
    if (num_tris) *num_tris = out.T;
    if (stats) *stats = out.st;
    return GCDT_OK;
    API_END
}

extern "C" int gcdt_solve_steiner(const gcdt_int2* points, int N, const gcdt_int2* cons, int C, int max_steiner,
                                  gcdt_int3* tris_out, gcdt_int3* adj_out, int* cflag_out, gcdt_hpoint* steiner_out,
                                  int* num_tris, int* num_steiner, gcdt_stats* stats, void* stream)
{
    API_BEGIN
    if (N < 0 || C < 0 || max_steiner < 0 || (N > 0 && !points) || (C > 0 && !cons) || !tris_out ||
        (max_steiner > 0 && !steiner_out))
        return fail(GCDT_ERR_INVALID, "bad arguments");
    if (N > (1 << 28) || (long long)N + max_steiner > (1 << 28)) return fail(GCDT_ERR_INVALID, "too many points");
    Ctx c((hipStream_t)stream);
    const long long tcap = 2LL * ((long long)N + max_steiner) + 2;
    DBuf<int3> adj_tmp;
    int3* adj = (int3*)adj_out;
    if (!adj) {
        adj_tmp.alloc((size_t)tcap);
        adj = adj_tmp.p;
    }
    SteinerArgs sa;
    sa.scap = max_steiner;
    sa.out = steiner_out;
    SolveOut out;
    solve_impl(c, (const int2*)points, N, (const int2*)cons, C, (int3*)tris_out, adj, out, cflag_out, &sa);
    if (num_tris) *num_tris = out.T;
    if (num_steiner) *num_steiner = out.S;
    if (stats) *stats = out.st;
    return GCDT_OK;
    API_END
}

// ------------------------------------------------------------------ reference entry points

namespace {

// Ctx on the legacy default stream with the caller's scratch attached.
struct ApiCtx {
    Ctx c;
    ApiCtx(void* tmp, size_t bytes) : c(nullptr)
    {
        if (tmp) {
            c.ext_scratch = (int*)tmp;
            c.ext_scratch_bytes = bytes;
        }
    }
};

__global__ void k_convert_sample(int3* tris, const int* sid, int ns, int np, int T)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        int3 t = tris[i];
        t.x = t.x >= ns ? t.x - ns + np : sid[t.x];
        t.y = t.y >= ns ? t.y - ns + np : sid[t.y];
        t.z = t.z >= ns ? t.z - ns + np : sid[t.z];
        tris[i] = t;
    }
}

__global__ void k_retire(const int* touched, const int* leaf_slot, int T, int* sons)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x)
        if (touched[i]) sons[leaf_slot[i]] = 0;
}

__global__ void k_leaf_slots(const int* sons, const int* off, int T, int* leaf_slot)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x)
        if (sons[i] == -1) leaf_slot[off[i]] = (int)i;
}

__global__ void k_flags_leaf(const int* sons, int T, int* f)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x)
        f[i] = sons[i] == -1;
}

__global__ void k_gather_touched(const int3* tri, const int* touched, const int* off, int T, int3* out)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x)
        if (touched[i]) out[off[i]] = tri[i];
}

// Insertion rounds from the two corner triangles over all points, with
// internal adjacency when the caller gives none.
static void raw_mesh_impl(Ctx& c, int2* pts, int3* tris, int3* adj_out, int* sons, int* numT, int n)
{
    long long cap = 4LL * n + 4;
    DBuf<int3> adj_tmp;
    int3* adj = adj_out;
    if (!adj) {
        adj_tmp.alloc(cap);
        adj = adj_tmp.p;
    }
    check_points(c, pts, n);
    init_square(c, pts, tris, adj, n);
    fill_i32(c, sons, -1, cap);
    DBuf<unsigned long long> claim(cap);
    fill_u64(c, claim.p, 0, cap);
    DBuf<int> loc((size_t)n + 4), pend(n > 0 ? n : 1), pend2(n > 0 ? n : 1);
    DBuf<unsigned long long> pkey((size_t)n + 4);
    locate_two(c, pts, n, loc.p, pend.p);
    Ins s;
    s.m = Mesh{pts, tris, adj, sons, nullptr};
    s.loc = loc.p;
    s.claim = claim.p;
    s.pkey = pkey.p;
    s.cap = (int)cap;
    s.err = c.dcnt.p + C_ERR;
    int r = 0;
    *numT = insert_rounds(c, s, pend, pend2, n, 2, &r);
}

}  // namespace

extern "C" {

int gcdt_get_sample_points(void* tmp, size_t tmpBytes, const gcdt_int2* points, int numPoints,
                           gcdt_int2** samplePoints, int** sample, int* num, double* ms)
{
    API_BEGIN
    if (numPoints < 0 || (numPoints > 0 && !points) || !samplePoints || !sample || !num)
        return fail(GCDT_ERR_INVALID, "bad arguments");
    ApiCtx a(tmp, tmpBytes);
    Timer t(a.c.st);
    check_points(a.c, (const int2*)points, numPoints);
    const size_t cap = (1 << 16) + 4;  // rawMesh.cu:715-717
    DBuf<int> ids(cap);
    DBuf<int2> sp(cap);
    int S = sample_points(a.c, (const int2*)points, numPoints, ids.p, sp.p);
    if (ms) *ms = t.ms();
    *num = S;
    *sample = ids.detach();
    *samplePoints = (gcdt_int2*)sp.detach();
    return GCDT_OK;
    API_END
}

int gcdt_raw_mesh_before(void* tmp, size_t tmpBytes, gcdt_int2* points, gcdt_int3* tris, int* sons,
                         int* numTriangles, int numPoints, double* ms)
{
    API_BEGIN
    if (numPoints < 0 || !points || !tris || !sons || !numTriangles) return fail(GCDT_ERR_INVALID, "bad arguments");
    ApiCtx a(tmp, tmpBytes);
    Timer t(a.c.st);
    raw_mesh_impl(a.c, (int2*)points, (int3*)tris, nullptr, sons, numTriangles, numPoints);
    if (ms) *ms = t.ms();
    return GCDT_OK;
    API_END
}

int gcdt_raw_mesh(void* tmp, size_t tmpBytes, gcdt_int2* points, gcdt_int3* tris, gcdt_int3* adjTris, int* sons,
                  int* numTriangles, int numPoints, double* ms)
{
    API_BEGIN
    if (numPoints < 0 || !points || !tris || !adjTris || !sons || !numTriangles)
        return fail(GCDT_ERR_INVALID, "bad arguments");
    ApiCtx a(tmp, tmpBytes);
    Timer t(a.c.st);
    raw_mesh_impl(a.c, (int2*)points, (int3*)tris, (int3*)adjTris, sons, numTriangles, numPoints);
    if (ms) *ms = t.ms();
    return GCDT_OK;
    API_END
}

int gcdt_raw_mesh_after(void* tmp, size_t tmpBytes, gcdt_int2* points, gcdt_int3* tris, gcdt_int3* adjTris,
                        int* sons, const int* bitmapOffsets, const int* bitmapTriIds, int* numTriangles,
                        int numPoints, double* ms)
{
    API_BEGIN
    if (numPoints < 0 || !points || !tris || !adjTris || !sons || !bitmapOffsets || !bitmapTriIds || !numTriangles ||
        *numTriangles < 2)
        return fail(GCDT_ERR_INVALID, "bad arguments");
    ApiCtx a(tmp, tmpBytes);
    Ctx& c = a.c;
    Timer t(c.st);
    int n = numPoints;
    long long cap = 4LL * n + 4;
    int T0 = *numTriangles;
    check_points(c, (const int2*)points, n);
    write_corners(c, (int2*)points, n);
    fill_i32(c, sons, -1, cap);
    DBuf<int> loc((size_t)n + 4), pend(n > 0 ? n : 1), pend2(n > 0 ? n : 1), isv(n > 0 ? n : 1);
    fill_i32(c, isv.p, 0, n);
    int np = locate_bitmap(c, (const int2*)points, n, (const int3*)tris, T0, bitmapOffsets, bitmapTriIds, loc.p,
                           isv.p, pend.p);
    DBuf<unsigned long long> claim(cap), pkey((size_t)n + 4);
    fill_u64(c, claim.p, 0, cap);
    Ins s;
    s.m = Mesh{(const int2*)points, (int3*)tris, (int3*)adjTris, sons, nullptr};
    s.loc = loc.p;
    s.claim = claim.p;
    s.pkey = pkey.p;
    s.cap = (int)cap;
    s.err = c.dcnt.p + C_ERR;
    int r = 0;
    *numTriangles = insert_rounds(c, s, pend, pend2, np, T0, &r);
    if (ms) *ms = t.ms();
    return GCDT_OK;
    API_END
}

int gcdt_get_bitmap(void* tmp, size_t tmpBytes, const gcdt_int2* points, const gcdt_int3* tris, int** bitmapOffsets,
                    int** bitmapTris, int numTriangles, double* ms)
{
    API_BEGIN
    if (numTriangles < 0 || !points || !tris || !bitmapOffsets || !bitmapTris) return fail(GCDT_ERR_INVALID, "bad arguments");
    ApiCtx a(tmp, tmpBytes);
    Timer t(a.c.st);
    int nb = 0;
    build_bitmap(a.c, (const int2*)points, (const int3*)tris, numTriangles, bitmapOffsets, bitmapTris, &nb);
    if (ms) *ms = t.ms();
    return GCDT_OK;
    API_END
}

int gcdt_filter_triangles(void* tmp, size_t tmpBytes, const gcdt_int3* tris, gcdt_int3* newTris,
                          const gcdt_int3* adjTris, gcdt_int3* newAdjTris, const int* sons, int* numTriangles,
                          double* ms)
{
    API_BEGIN
    if (!tris || !newTris || !sons || !numTriangles || *numTriangles < 0) return fail(GCDT_ERR_INVALID, "bad arguments");
    ApiCtx a(tmp, tmpBytes);
    Timer t(a.c.st);
    *numTriangles = filter_leaves(a.c, (const int3*)tris, (adjTris && newAdjTris) ? (const int3*)adjTris : nullptr,
                                  sons, *numTriangles, (int3*)newTris, newAdjTris ? (int3*)newAdjTris : nullptr,
                                  (long long)*numTriangles);
    if (ms) *ms = t.ms();
    return GCDT_OK;
    API_END
}

int gcdt_convert_sample(gcdt_int3* tris, const int* sampleId, int numSamples, int numPoints, int numTriangles)
{
    API_BEGIN
    if (!tris || !sampleId || numTriangles < 0) return fail(GCDT_ERR_INVALID, "bad arguments");
    if (numTriangles > 0) {
        k_convert_sample<<<grid_for(numTriangles), kBlock, 0, nullptr>>>((int3*)tris, sampleId, numSamples, numPoints,
                                                                         numTriangles);
        GCDT_LAUNCH_CHECK();
        GCDT_HIP(hipStreamSynchronize(nullptr));
    }
    return GCDT_OK;
    API_END
}

static int flip_plain(void* tmp, size_t tmpBytes, const gcdt_int2* points, gcdt_int3* tris, gcdt_int3* adjTris,
                      int numPoints, int numTriangles, double* ms)
{
    API_BEGIN
    if (!points || !tris || numTriangles < 0 || numPoints < 3) return fail(GCDT_ERR_INVALID, "bad arguments");
    ApiCtx a(tmp, tmpBytes);
    Ctx& c = a.c;
    Timer t(c.st);
    DBuf<int3> adj_tmp;
    int3* adj = (int3*)adjTris;
    if (!adj) {
        adj_tmp.alloc(numTriangles > 0 ? numTriangles : 1);
        adj = adj_tmp.p;
    }
    build_adjacency(c, (const int3*)tris, numTriangles, numPoints, adj);
    int r = 0;
    flip_rounds(c, Mesh{(const int2*)points, (int3*)tris, adj, nullptr, nullptr}, numTriangles, nullptr, 0, &r);
    if (ms) *ms = t.ms();
    return GCDT_OK;
    API_END
}

int gcdt_delaunay_sample(void* tmp, size_t tmpBytes, const gcdt_int2* points, gcdt_int3* tris, gcdt_int3* adjTris,
                         int numPoints, int numTriangles, double* ms)
{
    return flip_plain(tmp, tmpBytes, points, tris, adjTris, numPoints, numTriangles, ms);
}

int gcdt_try_fast_delaunay(void* tmp, size_t tmpBytes, const gcdt_int2* points, gcdt_int3* tris, gcdt_int3* adjTris,
                           int numPoints, int numTriangles, double* ms)
{
    return flip_plain(tmp, tmpBytes, points, tris, adjTris, numPoints, numTriangles, ms);
}

int gcdt_delaunay_new(void* tmp, size_t tmpBytes, gcdt_int2** points, gcdt_int3** tris, const gcdt_int2* cons,
                      int numPoints, int numTriangles, int numCons, double* ms)
{
    API_BEGIN
    if (!points || !*points || !tris || !*tris || numTriangles < 0 || numPoints < 3 || numCons < 0 ||
        (numCons > 0 && !cons))
        return fail(GCDT_ERR_INVALID, "bad arguments");
    ApiCtx a(tmp, tmpBytes);
    Ctx& c = a.c;
    Timer t(c.st);
    int T = numTriangles;
    DBuf<int3> adj(T > 0 ? T : 1);
    DBuf<int> cflag(T > 0 ? T : 1);
    fill_i32(c, cflag.p, 0, T);
    build_adjacency(c, (const int3*)*tris, T, numPoints, adj.p);
    Mesh m{(const int2*)*points, (int3*)*tris, adj.p, nullptr, cflag.p};
    mark_constraint_chains(c, m, T, numPoints, numPoints - 4, (const int2*)cons, numCons);
    int r = 0;
    flip_rounds(c, m, T, nullptr, 0, &r);
    if (ms) *ms = t.ms();
    return GCDT_OK;
    API_END
}

int gcdt_add_constraints(void* tmp, size_t tmpBytes, const gcdt_int2* points, gcdt_int3* tris, gcdt_int3* adjTris,
                         int* sons, const int* bitmapOffsets, const int* bitmapTris, const gcdt_int2* cons,
                         gcdt_int3** newTris, int numCons, int numPoints, int numTriangles, int* numNewTris, double* ms)
{
    API_BEGIN
    (void)bitmapOffsets;
    (void)bitmapTris;
    if (!points || !tris || !adjTris || !sons || !newTris || !numNewTris || numCons < 0 || numTriangles < 0 ||
        (numCons > 0 && !cons))
        return fail(GCDT_ERR_INVALID, "bad arguments");
    ApiCtx a(tmp, tmpBytes);
    Ctx& c = a.c;
    Timer t(c.st);
    int Tall = numTriangles;
    // compact the leaves (exact adjacency) and remember their slots
    DBuf<int> off((size_t)Tall + 1);
    k_flags_leaf<<<grid_for(Tall), kBlock, 0, c.st>>>(sons, Tall, off.p);
    GCDT_LAUNCH_CHECK();
    scan_excl_total(c, off.p, off.p, Tall);
    int T = c.get(C_TOTAL);
    DBuf<int3> tri(T > 0 ? T : 1), adj(T > 0 ? T : 1);
    DBuf<int> slot(T > 0 ? T : 1), cflag(T > 0 ? T : 1), touched(T > 0 ? T : 1);
    filter_leaves(c, (const int3*)tris, (const int3*)adjTris, sons, Tall, tri.p, adj.p, T);
    k_leaf_slots<<<grid_for(Tall), kBlock, 0, c.st>>>(sons, off.p, Tall, slot.p);
    GCDT_LAUNCH_CHECK();
    fill_i32(c, cflag.p, 0, T);
    fill_i32(c, touched.p, 0, T);
    Mesh m{(const int2*)points, tri.p, adj.p, nullptr, cflag.p};
    insert_constraints(c, m, T, numPoints + 4, numPoints, (const int2*)cons, numCons, touched.p, nullptr);
    // crossed triangles are retired in the DAG; their replacements go out
    k_retire<<<grid_for(T), kBlock, 0, c.st>>>(touched.p, slot.p, T, sons);
    GCDT_LAUNCH_CHECK();
    DBuf<int> toff((size_t)T + 1);
    GCDT_HIP(hipMemcpyAsync(toff.p, touched.p, sizeof(int) * (size_t)T, hipMemcpyDeviceToDevice, c.st));
    scan_excl_total(c, toff.p, toff.p, T);
    int nn = c.get(C_TOTAL);
    DBuf<int3> out(nn > 0 ? nn : 1);
    k_gather_touched<<<grid_for(T), kBlock, 0, c.st>>>(tri.p, touched.p, toff.p, T, out.p);
    GCDT_LAUNCH_CHECK();
    *numNewTris = nn;
    *newTris = (gcdt_int3*)out.detach();
    if (ms) *ms = t.ms();
    return GCDT_OK;
    API_END
}

}  // extern "C"

// ------------------------------------------------------------------ batches

namespace {

// One batched solve of K instances; poff/coff/toff (K+1 each) start at 0.
// Throws gcdt::Error on failure.
void batch_impl(const gcdt_int2* points, const std::vector<long long>& poff, const gcdt_int2* cons,
                const std::vector<long long>& coff, const std::vector<long long>& toff, int K, gcdt_int3* tris_out,
                gcdt_batch_stats& bs, hipStream_t st)
{
    long long NP = poff[K], NC = coff[K];
    long long nv = NP + 4LL * K;
    Ctx c(st);
    Timer total(c.st);
    DBuf<long long> d_poff(K + 1), d_coff(K + 1), d_toff(K + 1);
    GCDT_HIP(hipMemcpyAsync(d_poff.p, poff.data(), 8 * (K + 1), hipMemcpyHostToDevice, st));
    GCDT_HIP(hipMemcpyAsync(d_coff.p, coff.data(), 8 * (K + 1), hipMemcpyHostToDevice, st));
    GCDT_HIP(hipMemcpyAsync(d_toff.p, toff.data(), 8 * (K + 1), hipMemcpyHostToDevice, st));
    check_points(c, (const int2*)points, (int)NP);
    DBuf<int> vinst(NP > 0 ? NP : 1);
    batch_vinst(c, d_poff.p, K, NP, vinst.p);
    // points in (instance, Morton cell) order for locality, with about one cell
    // per point (4^L cells, L in [4, 10]); ids are mapped back at the end
    std::vector<int> lvl(K), cbase(K);
    long long ncells = 0;
    for (int i = 0; i < K; i++) {
        long long n = poff[i + 1] - poff[i];
        int L = 4;
        while (L < 10 && (1LL << (2 * (L + 1))) <= n) L++;
        lvl[i] = L;
        cbase[i] = (int)ncells;
        ncells += 1LL << (2 * L);
    }
    if (ncells > (1LL << 30)) throw Error(GCDT_ERR_INVALID, "batch too large");
    DBuf<int> d_lvl(K), d_cbase(K);
    GCDT_HIP(hipMemcpyAsync(d_lvl.p, lvl.data(), sizeof(int) * K, hipMemcpyHostToDevice, st));
    GCDT_HIP(hipMemcpyAsync(d_cbase.p, cbase.data(), sizeof(int) * K, hipMemcpyHostToDevice, st));
    DBuf<int2> pts(nv);
    DBuf<int> perm(NP > 0 ? NP : 1), inv(NP > 0 ? NP : 1), cell_end((size_t)ncells + 1);
    batch_order(c, (const int2*)points, vinst.p, d_cbase.p, d_lvl.p, NP, ncells, pts.p, perm.p, inv.p, cell_end.p);
    // ---- 1. sample meshes: one point per (instance, 16x16 cell), inserted from
    //         every instance's two seed triangles, then made Delaunay
    Timer t1(c.st);
    long long capA = 2LL * K + 4LL * std::min<long long>(NP, (long long)K << 8) + 8;
    DBuf<int> sid(std::max<long long>(1, std::min<long long>(NP, (long long)K << 8)));
    int S = batch_sample_sorted(c, d_cbase.p, d_lvl.p, cell_end.p, K, sid.p);
    cell_end.release();
    DBuf<int3> tA(capA), aA(capA);
    DBuf<int> sA(capA), loc(nv), pend(NP > 0 ? NP : 1), pend2(NP > 0 ? NP : 1);
    DBuf<unsigned long long> clA(capA), pkey(nv);
    fill_i32(c, sA.p, -1, capA);
    fill_u64(c, clA.p, 0, capA);
    batch_init(c, pts.p, tA.p, aA.p, K, NP);
    batch_locate_ids(c, pts.p, vinst.p, sid.p, S, loc.p, pend.p);
    sid.release();
    Ins s;
    s.m = Mesh{pts.p, tA.p, aA.p, sA.p, nullptr};
    s.loc = loc.p;
    s.claim = clA.p;
    s.pkey = pkey.p;
    s.cap = (int)capA;
    s.err = c.dcnt.p + C_ERR;
    int ri = 0;
    int ntA = insert_rounds(c, s, pend, pend2, S, 2 * K, &ri);
    long long capB = 2LL * K + 2LL * S + 4LL * (NP - S) + 8;
    if (capB > (1LL << 31) - 1) throw Error(GCDT_ERR_INVALID, "batch too large");
    DBuf<int3> tB(capB), aB(capB);
    int Ts = filter_leaves_inst(c, pts.p, vinst.p, d_cbase.p, d_lvl.p, ncells, NP, tA.p, aA.p, sA.p, ntA, tB.p,
                                aB.p, capB);
    tA.release();
    aA.release();
    sA.release();
    clA.release();
    validate_mesh(c, "the sample meshes' insertion", Mesh{pts.p, tB.p, aB.p, nullptr, nullptr}, nv, Ts, false);
    int fr = 0;
    bs.flips += flip_rounds(c, Mesh{pts.p, tB.p, aB.p, nullptr, nullptr}, Ts, nullptr, 0, &fr);
    validate_mesh(c, "the sample meshes' flips", Mesh{pts.p, tB.p, aB.p, nullptr, nullptr}, nv, Ts, true);
    // ---- 2. locate the other points through a per-instance bitmap of the
    //         sample meshes, then insert them
    int *bo = nullptr, *bt = nullptr, nbt = 0;
    build_bitmap_batch(c, pts.p, vinst.p, NP, K, tB.p, Ts, &bo, &bt, &nbt);
    DBuf<int> isv(NP > 0 ? NP : 1);
    fill_i32(c, isv.p, 0, NP);
    int np = locate_bitmap_batch(c, pts.p, vinst.p, NP, tB.p, Ts, bo, bt, loc.p, isv.p, pend.p);
    dev_give(bo, sizeof(int) * (size_t)(((long long)K << 8) + 1));
    dev_give(bt, sizeof(int) * (size_t)(nbt > 0 ? nbt : 1));
    isv.release();
    DBuf<int> sB(capB);
    DBuf<unsigned long long> clB(capB);
    fill_i32(c, sB.p, -1, capB);
    fill_u64(c, clB.p, 0, capB);
    Ins s2;
    s2.m = Mesh{pts.p, tB.p, aB.p, sB.p, nullptr};
    s2.loc = loc.p;
    s2.claim = clB.p;
    s2.pkey = pkey.p;
    s2.cap = (int)capB;
    s2.err = c.dcnt.p + C_ERR;
    int ri2 = 0;
    int nt = insert_rounds(c, s2, pend, pend2, np, Ts, &ri2);
    clB.release();
    pkey.release();
    loc.release();
    pend.release();
    pend2.release();
    long long Tn = 2 * NP + 2LL * K;
    DBuf<FRec> rec(Tn);
    int T = filter_leaves_inst_rec(c, pts.p, vinst.p, d_cbase.p, d_lvl.p, ncells, NP, tB.p, aB.p, sB.p, nt, rec.p,
                                   Tn);
    tB.release();
    aB.release();
    sB.release();
    DBuf<int3> tri(Tn), adj(Tn);  // after the insertion DAG is gone: a lower peak
    if (T != Tn) throw Error(GCDT_ERR_INTERNAL, "batch leaf count mismatch");
    bs.ms_insert = t1.ms();
    bs.insert_rounds = ri + ri2;
    // ---- Delaunay flips of all instances together (unpacked into tri / adj)
    Timer t2(c.st);
    int fr2 = 0;
    bs.flips += flip_rounds_records(c, rec.p, Mesh{pts.p, tri.p, adj.p, nullptr, nullptr}, T, &fr2);
    validate_mesh(c, "the flips", Mesh{pts.p, tri.p, adj.p, nullptr, nullptr}, nv, T, true);
    rec.release();
    bs.flip_rounds = fr + fr2;
    bs.ms_flip = t2.ms();
    // ---- constraints
    Timer t3(c.st);
    if (NC > 0) {
        DBuf<int2> cg(NC);
        batch_cons(c, (const int2*)cons, d_coff.p, d_poff.p, K, NC, cg.p);
        relabel_pairs(c, cg.p, (int)NC, inv.p, (int)NP, cg.p);
        DBuf<int> cflag(T), touched(T), tl(T);
        fill_i32(c, cflag.p, 0, T);
        fill_i32(c, touched.p, 0, T);
        Mesh m{pts.p, tri.p, adj.p, nullptr, cflag.p};
        ConsResult cr = insert_constraints(c, m, T, (int)nv, (int)NP, cg.p, (int)NC, touched.p, nullptr);
        bs.constraint_rounds = cr.rounds;
        bs.segments = cr.pieces;
        int ntl = flags_to_list(c, touched.p, T, tl.p);
        if (ntl > 0) {
            int f3 = 0;
            bs.flips += flip_rounds(c, m, T, tl.p, ntl, &f3);
            bs.flip_rounds += f3;
        }
        validate_mesh(c, "the constraints", m, nv, T, true);
    }
    bs.ms_constraints = t3.ms();
    batch_out(c, tri.p, T, vinst.p, perm.p, d_poff.p, d_toff.p, NP, K, (int3*)tris_out);
    bs.ms_total = total.ms();
    bs.num_instances = K;
    bs.num_points = NP;
    bs.num_triangles = T;
}

// statuses that come from an instance's own input (bad ids or coordinates,
// duplicate points, crossing constraints, and the consistency checks they
// can trip), as against the device, memory or the call's arguments
bool input_status(int s)
{
    return s == GCDT_ERR_INVALID || s == GCDT_ERR_NOT_CONVERGED || s == GCDT_ERR_CROSSING || s == GCDT_ERR_LOCATE ||
           s == GCDT_ERR_CAPACITY || s == GCDT_ERR_INTERNAL;
}

// Instances [a, b) as one batch; when it fails with an input status, its two
// halves are solved separately, down to single instances, whose status is
// recorded.  Each instance's triangles land in its own range of tris_out.
void batch_isolate(const gcdt_int2* points, const std::vector<long long>& poff, const gcdt_int2* cons,
                   const std::vector<long long>& coff, const std::vector<long long>& toff, int a, int b,
                   gcdt_int3* tris_out, gcdt_batch_stats& acc, int* inst_status, std::string& first_err,
                   hipStream_t st)
{
    std::vector<long long> p(b - a + 1), c(b - a + 1), t(b - a + 1);
    for (int i = a; i <= b; i++) {
        p[i - a] = poff[i] - poff[a];
        c[i - a] = coff[i] - coff[a];
        t[i - a] = toff[i] - toff[a];
    }
    gcdt_batch_stats bs{};
    try {
        batch_impl(points + poff[a], p, cons ? cons + coff[a] : nullptr, c, t, b - a, tris_out + toff[a], bs, st);
    } catch (const Error& e) {
        if (!inst_status || !input_status(e.status)) throw;
        if (st) (void)hipStreamSynchronize(st);
        if (b - a == 1) {
            inst_status[a] = e.status;
            if (first_err.empty()) first_err = "instance " + std::to_string(a) + ": " + e.what();
            return;
        }
        int m = a + (b - a) / 2;
        batch_isolate(points, poff, cons, coff, toff, a, m, tris_out, acc, inst_status, first_err, st);
        batch_isolate(points, poff, cons, coff, toff, m, b, tris_out, acc, inst_status, first_err, st);
        return;
    }
    for (int i = a; i < b; i++)
        if (inst_status) inst_status[i] = GCDT_OK;
    acc.ms_total += bs.ms_total;
    acc.ms_insert += bs.ms_insert;
    acc.ms_flip += bs.ms_flip;
    acc.ms_constraints += bs.ms_constraints;
    acc.insert_rounds = std::max(acc.insert_rounds, bs.insert_rounds);
    acc.flip_rounds = std::max(acc.flip_rounds, bs.flip_rounds);
    acc.constraint_rounds = std::max(acc.constraint_rounds, bs.constraint_rounds);
    acc.num_instances += bs.num_instances;
    acc.num_points += bs.num_points;
    acc.num_triangles += bs.num_triangles;
    acc.flips += bs.flips;
    acc.segments += bs.segments;
}

}  // namespace

extern "C" int gcdt_solve_batch(const gcdt_int2* points, const int64_t* point_offsets, int K, const gcdt_int2* cons,
                                const int64_t* cons_offsets, gcdt_int3* tris_out, gcdt_batch_stats* stats,
                                int* inst_status, void* stream)
{
    API_BEGIN
    if (K <= 0 || !point_offsets || !tris_out) return fail(GCDT_ERR_INVALID, "bad arguments");
    if (point_offsets[0] != 0) return fail(GCDT_ERR_INVALID, "point_offsets[0] must be 0");
    std::vector<long long> poff(K + 1), coff(K + 1, 0), toff(K + 1, 0);
    for (int i = 0; i <= K; i++) {
        poff[i] = point_offsets[i];
        if (i && poff[i] < poff[i - 1]) return fail(GCDT_ERR_INVALID, "point_offsets not monotone");
        if (cons_offsets) {
            coff[i] = cons_offsets[i];
            if (i && coff[i] < coff[i - 1]) return fail(GCDT_ERR_INVALID, "cons_offsets not monotone");
        }
        if (i) toff[i] = toff[i - 1] + 2 * (poff[i] - poff[i - 1]) + 2;
    }
    if (cons_offsets && coff[0] != 0) return fail(GCDT_ERR_INVALID, "cons_offsets[0] must be 0");
    long long NP = poff[K], NC = coff[K];
    if (NP > 0 && !points) return fail(GCDT_ERR_INVALID, "no points");
    if (NC > 0 && !cons) return fail(GCDT_ERR_INVALID, "no constraints");
    long long nv = NP + 4LL * K, cap = 2LL * K + 4 * NP + 8;
    if (cap > (1LL << 30) || nv > (1LL << 29)) return fail(GCDT_ERR_INVALID, "batch too large");
    // the per-instance location grid has 256 cells per instance
    if (((long long)K << 8) > (1LL << 26)) return fail(GCDT_ERR_INVALID, "more than 262144 instances in one batch");
    hipStream_t st = (hipStream_t)stream;
    gcdt_batch_stats bs{};
    std::string first_err;
    batch_isolate(points, poff, cons, coff, toff, 0, K, tris_out, bs, inst_status, first_err, st);
    bs.num_instances = K;
    if (stats) *stats = bs;
    if (!first_err.empty()) g_last_error = first_err;
    return GCDT_OK;
    API_END
}

extern "C" int gcdt_set_tunable(const char* name, int value)
{
    if (!name) return fail(GCDT_ERR_INVALID, "null name");
    std::string n(name);
    auto range = [&](int lo, int hi) {
        return value >= lo && value <= hi ? GCDT_OK
                                          : fail(GCDT_ERR_INVALID, n + " must lie in [" + std::to_string(lo) + ", " +
                                                                       std::to_string(hi) + "]");
    };
    int rc;
    // only values that a GPU parity test covers are accepted
    auto cleared = [&]() {
        return tunable_cleared(n, value)
                   ? GCDT_OK
                   : fail(GCDT_ERR_INVALID, n + "=" + std::to_string(value) + " is not covered by a GPU parity test");
    };
    if (n == "tail_threshold") {
        if ((rc = range(0, 1 << 20)) || (rc = cleared())) return rc;
        Tunables::get().tail_threshold = value;
    } else if (n == "pivot") {
        if ((rc = range(0, 1)) || (rc = cleared())) return rc;
        Tunables::get().pivot = value;
    } else if (n == "single_tester") {
        if ((rc = range(0, 1)) || (rc = cleared())) return rc;
        Tunables::get().single_tester = value;
    } else if (n == "validate") {
        if ((rc = range(0, 1)) || (rc = cleared())) return rc;
        Tunables::get().validate = value;
    } else if (n == "grid_cap") {
        if ((rc = range(256, 65536)) || (rc = cleared())) return rc;
        g_grid_cap = value;
    } else if (n == "flip_items" || n == "ins_items") {
        if (value != 1 && value != 2 && value != 4) return fail(GCDT_ERR_INVALID, n + " must be 1, 2 or 4");
        if ((rc = cleared())) return rc;
        (n == "flip_items" ? Tunables::get().flip_items : Tunables::get().ins_items) = value;
    } else {
        return fail(GCDT_ERR_INVALID, "unknown tunable " + n);
    }
    return GCDT_OK;
}

extern "C" int gcdt_get_tunable(const char* name, int* value)
{
    if (!name || !value) return fail(GCDT_ERR_INVALID, "null argument");
    std::string n(name);
    if (n == "tail_threshold") *value = Tunables::get().tail_threshold;
    else if (n == "pivot") *value = Tunables::get().pivot;
    else if (n == "single_tester") *value = Tunables::get().single_tester;
    else if (n == "validate") *value = Tunables::get().validate;
    else if (n == "grid_cap") *value = g_grid_cap;
    else if (n == "flip_items") *value = Tunables::get().flip_items;
    else if (n == "ins_items") *value = Tunables::get().ins_items;
    else return fail(GCDT_ERR_INVALID, "unknown tunable " + n);
    return GCDT_OK;
}

extern "C" int gcdt_check_mesh(const gcdt_int2* points, long long nv, const gcdt_int3* tris, const gcdt_int3* adj,
                               const int* cflag, int T, int delaunay, int* failed, int* first_bad, void* stream)
{
    API_BEGIN
    if (!points || nv <= 0 || T < 0 || (T > 0 && (!tris || !adj)) || !failed || !first_bad)
        return fail(GCDT_ERR_INVALID, "bad arguments");
    Ctx c((hipStream_t)stream);
    check_mesh(c, (const int2*)points, nv, (const int3*)tris, (const int3*)adj, cflag, T, delaunay != 0, failed,
               first_bad);
    return GCDT_OK;
    API_END
}
