// gfx950 kernels and host drivers of the mesh phases: scans, sampling,
// bitmap ("boostmap") location, parallel insertion rounds, flip rounds,
// leaf compaction and adjacency construction.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "mesh.hpp"

namespace gcdt {

// ================================================================== context

// Values of each tunable that a GPU parity test covers (the autotuner's
// candidates, plus tail thresholds 0 and 2^20 that force one round path).
// Anything else is refused, by gcdt_set_tunable and by the environment
// overrides alike.
bool tunable_cleared(const std::string& name, int v)
{
    if (name == "tail_threshold")
        return v == 0 || v == 1024 || v == 2048 || v == 4096 || v == 8192 || v == 16384 || v == (1 << 20);
    if (name == "grid_cap") return v == 2048 || v == 4096 || v == 8192 || v == 16384;
    if (name == "flip_items" || name == "ins_items") return v == 1 || v == 2 || v == 4;
    if (name == "pivot" || name == "single_tester" || name == "validate") return v == 0 || v == 1;
    return false;
}

Tunables& Tunables::get()
{
    static Tunables t = [] {
        Tunables x;
        auto env = [&](const char* var, const char* name, int* dst) {
            const char* e = getenv(var);
            if (!e) return;
            char* end = nullptr;
            long v = strtol(e, &end, 10);
            if (end == e || *end != 0 || v < INT_MIN || v > INT_MAX || !tunable_cleared(name, (int)v)) {
                x.env_error += std::string(var) + "=" + e + " is not a cleared value; ";
                return;
            }
            *dst = (int)v;
        };
        env("GCDT_TAIL_THRESHOLD", "tail_threshold", &x.tail_threshold);
        env("GCDT_PIVOT", "pivot", &x.pivot);
        env("GCDT_SINGLE_TESTER", "single_tester", &x.single_tester);
        env("GCDT_VALIDATE", "validate", &x.validate);
        env("GCDT_GRID_CAP", "grid_cap", &g_grid_cap);
        return x;
    }();
    return t;
}

thread_local std::vector<PendingFree>* g_defer = nullptr;
int g_grid_cap = kMaxGrid;

thread_local Arena* g_arena = nullptr;

void* Arena::take(size_t bytes)
{
    size_t len = (bytes + 255) & ~(size_t)255;
    if (len > cap - top)
        throw Error(GCDT_ERR_SCRATCH, "the scratch array (" + std::to_string(cap) + " bytes) is too small for this "
                                      "input: " + std::to_string(top + len) + " bytes needed so far");
    void* p = base + top;
    top += len;
    if (top > high) high = top;
    return p;
}

void Arena::give(void* p, size_t bytes)
{
    size_t off = (size_t)((char*)p - base), end = off + ((bytes + 255) & ~(size_t)255);
    if (end != top) {
        if (nfreed < kFreed) {
            freed_off[nfreed] = off;
            freed_end[nfreed] = end;
            nfreed++;
        }
        return;  // a full list leaves the block in use until the call ends
    }
    top = off;
    for (bool again = true; again;) {
        again = false;
        for (int i = 0; i < nfreed; i++)
            if (freed_end[i] == top) {
                top = freed_off[i];
                freed_off[i] = freed_off[--nfreed];
                freed_end[i] = freed_end[nfreed];
                again = true;
                break;
            }
    }
}

void* dev_take(size_t bytes)
{
    if (g_arena) return g_arena->take(bytes);
    void* p = getMemCacheRef().getMemByAbtrtSize(bytes);
    if (!p) throw Error(GCDT_ERR_OOM, "device pool allocation failed");
    return p;
}

void dev_give(void* p, size_t bytes)
{
    if (!p) return;
    if (g_arena && g_arena->owns(p)) g_arena->give(p, bytes);
    else if (g_defer) g_defer->push_back({p, bytes});
    else getMemCacheRef().releaseAbtrtMem(p, bytes);
}

namespace {
// Pinned counter mirrors and snapshot rings, reused across solves: allocating
// and freeing pinned memory per call costs more than a small solve and
// hipHostFree waits for the whole device, which would serialise solves that
// run on different streams.
std::mutex g_host_mu;
std::vector<HostSlot> g_host_free;
}  // namespace

HostSlot take_host_slot()
{
    {
        std::lock_guard<std::mutex> g(g_host_mu);
        if (!g_host_free.empty()) {
            HostSlot h = g_host_free.back();
            g_host_free.pop_back();
            return h;
        }
    }
    HostSlot h{};
    GCDT_HIP(hipHostMalloc((void**)&h.hcnt, sizeof(int) * C_NUM, hipHostMallocDefault));
    // written by kernels (round-begin snapshots): coherent, so the writes reach
    // host memory without a cache flush; sequence number 0 is never issued
    GCDT_HIP(hipHostMalloc((void**)&h.hring, sizeof(unsigned long long) * 2 * C_NUM,
                           hipHostMallocMapped | hipHostMallocCoherent));
    memset(h.hring, 0, sizeof(unsigned long long) * 2 * C_NUM);
    GCDT_HIP(hipHostGetDevicePointer((void**)&h.hring_dev, h.hring, 0));
    return h;
}

Ctx::Ctx(hipStream_t s, const HostSlot* fixed) : st(s)
{
    if (!Tunables::get().env_error.empty()) throw Error(GCDT_ERR_INVALID, Tunables::get().env_error);
    tail_threshold = Tunables::get().tail_threshold;
    pivot = Tunables::get().pivot;
    single_tester = Tunables::get().single_tester;
    flip_items = Tunables::get().flip_items;
    ins_items = Tunables::get().ins_items;
    dcnt.alloc(C_NUM);
    HostSlot h = fixed ? *fixed : take_host_slot();
    fixed_slot = fixed != nullptr;
    hcnt = h.hcnt;
    hring = h.hring;
    hring_dev = h.hring_dev;
    if (st) {
        prev_defer = g_defer;
        g_defer = &deferred;
    }
    zero_counters();
}

Ctx::~Ctx()
{
    if (st) {
        // every kernel of this solve has finished before its blocks go back
        (void)hipStreamSynchronize(st);
        g_defer = prev_defer;
        for (auto& f : deferred) getMemCacheRef().releaseAbtrtMem(f.p, f.bytes);
        deferred.clear();
    }
    if (hcnt && !fixed_slot) {
        std::lock_guard<std::mutex> g(g_host_mu);
        g_host_free.push_back(HostSlot{hcnt, hring, hring_dev});
    }
}

// One counter into a snapshot slot: one 64-bit store at system scope, so the
// host never sees a value without its sequence number.
__device__ inline void snap_store(unsigned long long* dst, unsigned seq, int v)
{
    __hip_atomic_store(dst, ((unsigned long long)seq << 32) | (unsigned)v, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

// the counters straight into the mapped pinned ring (a kernel costs less
// than a copy-engine transfer)
__global__ void k_snap(const int* cnt, unsigned long long* dst, unsigned seq)
{
    snap_store(&dst[threadIdx.x], seq, cnt[threadIdx.x]);
}

namespace {
// process-wide, so a slot reused by another context never matches old words
std::atomic<unsigned> g_snap_seq{0};
}  // namespace

unsigned Ctx::snap_begin(int slot)
{
    unsigned s = ++g_snap_seq;
    if (s == 0) s = ++g_snap_seq;
    ring_seq[slot] = s;
    return s;
}

void Ctx::snapshot(int slot)
{
    unsigned seq = snap_begin(slot);
    k_snap<<<1, C_NUM, 0, st>>>(dcnt.p, snap_dst(slot), seq);
    GCDT_LAUNCH_CHECK();
}

const int* Ctx::wait_snapshot(int slot)
{
    const unsigned long long* p = hring + C_NUM * slot;
    const unsigned want = ring_seq[slot];
    int* out = hsnap[slot];
    auto complete = [&]() {
        for (int i = 0; i < C_NUM; i++) {
            unsigned long long v = __atomic_load_n(p + i, __ATOMIC_ACQUIRE);
            if ((unsigned)(v >> 32) != want) return false;
            out[i] = (int)(unsigned)v;
        }
        return true;
    };
    for (unsigned spin = 1;; ++spin) {
        if (complete()) return out;
        if ((spin & 1023) == 0) {
            // the words can be missing only while their kernel is queued or
            // running: a device error or a drained stream without them ends
            // the wait
            hipError_t e = hipStreamQuery(st);
            if (e == hipSuccess) {
                if (complete()) return out;
                throw Error(GCDT_ERR_INTERNAL, "counter snapshot missing");
            }
            if (e != hipErrorNotReady) GCDT_HIP(e);
        }
    }
}

void Ctx::read(int i, int k)
{
    GCDT_HIP(hipMemcpyAsync(hcnt + i, dcnt.p + i, sizeof(int) * k, hipMemcpyDeviceToHost, st));
    GCDT_HIP(hipStreamSynchronize(st));
}

__global__ void k_set1(int* p, int v) { *p = v; }

void Ctx::set(int i, int v)
{
    k_set1<<<1, 1, 0, st>>>(dcnt.p + i, v);
    GCDT_LAUNCH_CHECK();
}

void Ctx::zero_counters() { fill_i32(*this, dcnt.p, 0, C_NUM); }

void Ctx::check_err()
{
    int e = get(C_ERR);
    if (e == 0) return;
    int status = GCDT_ERR_INTERNAL;
    if (e & MERR_CAPACITY) status = GCDT_ERR_CAPACITY;
    else if (e & (MERR_LOCATE | MERR_DUPLICATE)) status = GCDT_ERR_LOCATE;
    else if (e & MERR_CROSSING) status = GCDT_ERR_CROSSING;
    else if (e & MERR_SCRATCH) status = GCDT_ERR_SCRATCH;
    else if (e & MERR_INVALID) status = GCDT_ERR_INVALID;
    throw Error(status, "device consistency flag " + std::to_string(e));
}

// ================================================================== live kernel timing

namespace {
struct ProfRec {
    std::string name;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    long long launches = 0;
    long long items = 0;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::string g_prof_only;            // time only this kernel when not empty
std::vector<ProfRec> g_prof;
std::vector<hipEvent_t> g_ev_pool;  // events are reused across resets

bool take_event(hipEvent_t* e)
{
    if (!g_ev_pool.empty()) {
        *e = g_ev_pool.back();
        g_ev_pool.pop_back();
        return true;
    }
    return hipEventCreate(e) == hipSuccess;
}
}  // namespace

ProfScope::ProfScope(const char* name, hipStream_t s, long long items) : st(s)
{
    // no events (and no bookkeeping that allocates) inside an arena call
    if (!g_prof_on || g_arena) return;
    std::lock_guard<std::mutex> g(g_prof_mu);
    if (!g_prof_only.empty() && g_prof_only != name) return;
    for (size_t i = 0; i < g_prof.size(); i++)
        if (g_prof[i].name == name) slot = (int)i;
    if (slot < 0) {
        g_prof.push_back(ProfRec{name, {}, 0, 0});
        slot = (int)g_prof.size() - 1;
    }
    hipEvent_t a, b;
    if (!take_event(&a)) {
        slot = -1;
        return;
    }
    if (!take_event(&b)) {
        g_ev_pool.push_back(a);
        slot = -1;
        return;
    }
    (void)hipEventRecord(a, st);
    g_prof[slot].ev.push_back({a, b});
    g_prof[slot].launches++;
    g_prof[slot].items += items;
}

ProfScope::~ProfScope()
{
    if (slot < 0) return;
    std::lock_guard<std::mutex> g(g_prof_mu);
    (void)hipEventRecord(g_prof[slot].ev.back().second, st);
}

void prof_enable(bool on)
{
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_prof_on = on;
}

void prof_reset()
{
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (auto& r : g_prof)
        for (auto& e : r.ev) {
            (void)hipEventSynchronize(e.second);
            g_ev_pool.push_back(e.first);
            g_ev_pool.push_back(e.second);
        }
    g_prof.clear();
}

void prof_add_items(const char* name, long long items)
{
    if (!g_prof_on) return;
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (auto& r : g_prof)
        if (r.name == name) {
            r.items += items;
            return;
        }
    // an item count without timed launches of its own (work done inside
    // another kernel's launches)
    g_prof.push_back(ProfRec{name, {}, 0, items});
}

void prof_only(const char* name)
{
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_prof_only = name ? name : "";
}

// total ms, launches and items of one kernel (synchronises its events)
bool prof_read(const char* name, double* ms, long long* launches, long long* items)
{
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (auto& r : g_prof) {
        if (r.name != name) continue;
        double t = 0;
        for (auto& e : r.ev) {
            (void)hipEventSynchronize(e.second);
            float f = 0;
            (void)hipEventElapsedTime(&f, e.first, e.second);
            t += f;
        }
        *ms = t;
        *launches = r.launches;
        *items = r.items;
        return true;
    }
    return false;
}

int prof_names(char* buf, int cap)
{
    std::lock_guard<std::mutex> g(g_prof_mu);
    std::string all;
    for (auto& r : g_prof) {
        all += r.name;
        all += ';';
    }
    if (buf && cap > 0) {
        int n = (int)std::min<size_t>(all.size(), (size_t)cap - 1);
        memcpy(buf, all.data(), n);
        buf[n] = 0;
    }
    return (int)all.size();
}

// ================================================================== primitives

__global__ __launch_bounds__(kBlock) void k_fill_i32(int* p, int v, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_fill_u64(unsigned long long* p, unsigned long long v, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_i32(Ctx& c, int* p, int v, long long n)
{
    if (n <= 0) return;
    k_fill_i32<<<grid_for(n), kBlock, 0, c.st>>>(p, v, n);
    GCDT_LAUNCH_CHECK();
}

void fill_u64(Ctx& c, unsigned long long* p, unsigned long long v, long long n)
{
    if (n <= 0) return;
    k_fill_u64<<<grid_for(n), kBlock, 0, c.st>>>(p, v, n);
    GCDT_LAUNCH_CHECK();
}

constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = kBlock * SCAN_ITEMS;
constexpr int SCAN_PAD(int i) { return i + (i >> 4); }

__device__ inline int block_excl_scan(int v, int& total)
{
    __shared__ int wsum[kBlock / 64];
    int lane = __lane_id(), w = threadIdx.x >> 6;
    int wt;
    int ex = wave_excl_scan(v, wt);
    if (lane == 0) wsum[w] = wt;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; i++) {
        int s = wsum[i];
        if (i < w) base += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return base + ex;
}

__global__ __launch_bounds__(kBlock) void k_scan_tile_sums(const int* in, long long n, int* sums)
{
    long long base = (long long)blockIdx.x * SCAN_TILE;
    int s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        long long i = base + (long long)j * kBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    int tot;
    block_excl_scan(s, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_sums(int* sums, int nt, int* total, int* total2)
{
    __shared__ int wsum[16];
    __shared__ int carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int base = 0; base < nt; base += 1024) {
        int i = base + threadIdx.x;
        int v = i < nt ? sums[i] : 0;
        int lane = __lane_id(), w = threadIdx.x >> 6;
        int wt;
        int ex = wave_excl_scan(v, wt);
        if (lane == 0) wsum[w] = wt;
        __syncthreads();
        int pre = 0, tot = 0;
        for (int q = 0; q < 16; q++) {
            if (q < w) pre += wsum[q];
            tot += wsum[q];
        }
        int carry = carry_s;
        if (i < nt) sums[i] = carry + pre + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry_s = carry + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (total) *total = carry_s;
        if (total2) *total2 = carry_s;
    }
}

template <bool INCL>
__global__ __launch_bounds__(kBlock) void k_scan_apply(const int* in, int* out, long long n, const int* sums)
{
    __shared__ int tile[SCAN_PAD(SCAN_TILE)];
    long long base = (long long)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        int li = j * kBlock + threadIdx.x;
        long long i = base + li;
        tile[SCAN_PAD(li)] = i < n ? in[i] : 0;
    }
    __syncthreads();
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        v[j] = tile[SCAN_PAD(threadIdx.x * SCAN_ITEMS + j)];
        s += v[j];
    }
    int tot;
    int run = sums[blockIdx.x] + block_excl_scan(s, tot);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        run += INCL ? v[j] : 0;
        tile[SCAN_PAD(threadIdx.x * SCAN_ITEMS + j)] = run;
        run += INCL ? 0 : v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        int li = j * kBlock + threadIdx.x;
        long long i = base + li;
        if (i < n) out[i] = tile[SCAN_PAD(li)];
    }
}

long long scan_scratch_ints(long long n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

// Exclusive (inclusive when incl) scan of in[0..n) into out[0..n); the total
// goes to dcnt[C_TOTAL] and, when total_at_end, also to out[n] (out must then
// hold n+1 ints).
static void scan_impl(Ctx& c, const int* in, int* out, long long n, bool total_at_end, bool incl = false)
{
    if (n <= 0) {
        c.set(C_TOTAL, 0);
        if (total_at_end) {
            k_set1<<<1, 1, 0, c.st>>>(out, 0);
            GCDT_LAUNCH_CHECK();
        }
        return;
    }
    long long nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    int* sums;
    if (c.ext_scratch) {
        if (c.ext_scratch_bytes < sizeof(int) * (size_t)(nt + 1))
            throw Error(GCDT_ERR_SCRATCH, "caller scratch too small for a scan of " + std::to_string(n));
        sums = c.ext_scratch;
    } else {
        c.scan_sums.ensure((size_t)nt + 1);
        sums = c.scan_sums.p;
    }
    k_scan_tile_sums<<<(unsigned)nt, kBlock, 0, c.st>>>(in, n, sums);
    GCDT_LAUNCH_CHECK();
    k_scan_sums<<<1, 1024, 0, c.st>>>(sums, (int)nt, c.dcnt.p + C_TOTAL, total_at_end ? out + n : nullptr);
    GCDT_LAUNCH_CHECK();
    if (incl)
        k_scan_apply<true><<<(unsigned)nt, kBlock, 0, c.st>>>(in, out, n, sums);
    else
        k_scan_apply<false><<<(unsigned)nt, kBlock, 0, c.st>>>(in, out, n, sums);
    GCDT_LAUNCH_CHECK();
}

void scan_excl(Ctx& c, const int* in, int* out, long long n) { scan_impl(c, in, out, n, false); }
void scan_incl(Ctx& c, const int* in, int* out, long long n) { scan_impl(c, in, out, n, false, true); }
void scan_excl_total(Ctx& c, const int* in, int* out, long long n) { scan_impl(c, in, out, n, true); }

// ================================================================== square / sample

constexpr int GRID_R = (1 << 30) - 1;

__global__ void k_init_square(int2* pts, int3* tri, int3* adj, int n)
{
    if (pts) {
        pts[n] = make_int2(0, 0);
        pts[n + 1] = make_int2(GRID_R, 0);
        pts[n + 2] = make_int2(GRID_R, GRID_R);
        pts[n + 3] = make_int2(0, GRID_R);
    }
    // rawMesh.cu:16-21: (N,N+1,N+3) and (N+1,N+2,N+3) split along the anti-diagonal
    if (tri) {
        tri[0] = make_int3(n, n + 1, n + 3);
        tri[1] = make_int3(n + 1, n + 2, n + 3);
    }
    if (adj) {
        adj[0] = make_int3(1, -1, -1);
        adj[1] = make_int3(-1, 0, -1);
    }
}

__global__ __launch_bounds__(kBlock) void k_check_points(const int2* pts, int n, int* err)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        int2 p = pts[i];
        if (p.x <= 0 || p.y <= 0 || p.x >= GRID_R || p.y >= GRID_R) atomicOr(err, MERR_INVALID);
    }
}

void check_points(Ctx& c, const int2* pts, int n)
{
    if (n <= 0) return;
    k_check_points<<<grid_for(n), kBlock, 0, c.st>>>(pts, n, c.dcnt.p + C_ERR);
    GCDT_LAUNCH_CHECK();
    c.check_err();
}

void init_square(Ctx& c, int2* pts, int3* tri, int3* adj, int n)
{
    k_init_square<<<1, 1, 0, c.st>>>(pts, tri, adj, n);
    GCDT_LAUNCH_CHECK();
}

__global__ void k_write_corners(int2* pts, int n)
{
    pts[n] = make_int2(0, 0);
    pts[n + 1] = make_int2(GRID_R, 0);
    pts[n + 2] = make_int2(GRID_R, GRID_R);
    pts[n + 3] = make_int2(0, GRID_R);
}

void write_corners(Ctx& c, int2* pts, int n)
{
    k_write_corners<<<1, 1, 0, c.st>>>(pts, n);
    GCDT_LAUNCH_CHECK();
}

constexpr int LOG_SAMPLE = 8;   // math_base.cuh:24 logSample
constexpr int LOG_BITMAP = 10;  // math_base.cuh:25 logBitmap
constexpr int SAMPLE_CELLS = 1 << (2 * LOG_SAMPLE);
constexpr int BITMAP_CELLS = 1 << (2 * LOG_BITMAP);

// Cell grids of the point-location bitmap ("boostmap").  GridOne: one mesh,
// 1024x1024 cells over the square (math_base.cuh:25 logBitmap).  GridInst:
// a batch of instances that all span the square, each with its own 16x16
// cells; a cell id is instance * 256 + local cell.
struct GridOne {
    static constexpr int LOG = LOG_BITMAP;
    __device__ long long tri_base(const int3&) const { return 0; }
    __device__ long long pt_base(int) const { return 0; }
};

struct GridInst {
    static constexpr int LOG = 4;
    const int* vinst;
    long long NP;  // corners of instance i are NP + 4i .. NP + 4i + 3
    __device__ int inst_of(int v) const { return v < NP ? vinst[v] : (int)((v - NP) >> 2); }
    __device__ long long tri_base(const int3& t) const { return (long long)inst_of(t.x) << (2 * LOG); }
    __device__ long long pt_base(int p) const { return (long long)vinst[p] << (2 * LOG); }
};

template <class G>
__device__ inline long long pt_cell(const G& g, int p, int2 P)
{
    return g.pt_base(p) + ((long long)(P.x >> (30 - G::LOG)) << G::LOG) + (P.y >> (30 - G::LOG));
}

__global__ __launch_bounds__(kBlock) void k_sample_cells(const int2* pts, int n, int* cellmin)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        int2 p = pts[i];
        int cell = ((p.x >> (30 - LOG_SAMPLE)) << LOG_SAMPLE) + (p.y >> (30 - LOG_SAMPLE));
        atomicMin(&cellmin[cell], (int)i);
    }
}

__global__ __launch_bounds__(kBlock) void k_sample_flags(const int* cellmin, int* flags)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < SAMPLE_CELLS; i += gridDim.x * blockDim.x)
        flags[i] = cellmin[i] != INT_MAX;
}

__global__ __launch_bounds__(kBlock) void k_sample_emit(const int2* pts, const int* cellmin, const int* offs, int* ids,
                                                        int2* spts)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < SAMPLE_CELLS; i += gridDim.x * blockDim.x) {
        int p = cellmin[i];
        if (p != INT_MAX) {
            int o = offs[i];
            if (ids) ids[o] = p;
            if (spts) spts[o] = pts[p];
        }
    }
}

int sample_points(Ctx& c, const int2* pts, int n, int* sample_ids, int2* sample_pts)
{
    DBuf<int> cellmin(SAMPLE_CELLS), offs(SAMPLE_CELLS + 1);
    fill_i32(c, cellmin.p, INT_MAX, SAMPLE_CELLS);
    if (n > 0) {
        k_sample_cells<<<grid_for(n), kBlock, 0, c.st>>>(pts, n, cellmin.p);
        GCDT_LAUNCH_CHECK();
    }
    k_sample_flags<<<grid_for(SAMPLE_CELLS), kBlock, 0, c.st>>>(cellmin.p, offs.p);
    GCDT_LAUNCH_CHECK();
    scan_excl(c, offs.p, offs.p, SAMPLE_CELLS);
    k_sample_emit<<<grid_for(SAMPLE_CELLS), kBlock, 0, c.st>>>(pts, cellmin.p, offs.p, sample_ids, sample_pts);
    GCDT_LAUNCH_CHECK();
    return c.get(C_TOTAL);
}

// ================================================================== Morton order

// 20-bit Morton code of the 1024x1024 cell of a 2^30-grid point
__device__ inline unsigned morton_cell(int2 p)
{
    unsigned x = (unsigned)p.x >> 20, y = (unsigned)p.y >> 20;
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

// ---------------------------------------------------------------- order inside cells

// Counting sorts place the elements of one bucket in the order their atomics
// arrived.  Where that order reaches a decision (point ids key the insertion
// claims, triangle ids key the flip claims), each bucket's range is then
// reordered by a key of its own elements -- the input index of a point, the
// sorted vertex ids of a triangle -- so the order is the same on every run.
// The scatter writes one slot per position: (element, bucket) for triangles,
// (element, bucket, x, y) for points.  A bucket of up to CELL_RANK_MAX (256)
// elements is ranked by one lane per element, its key
// against every key of the bucket (all staged in LDS by the workgroup); buckets
// of up to CELL_BLOCK_MAX by one workgroup's bitonic sort in LDS; larger ones,
// which the count passes flag, by a merge sort of all their elements
// (cell_reorder_huge).
#ifndef GCDT_CELL_RANK_MAX
#define GCDT_CELL_RANK_MAX 256
#endif
#ifndef GCDT_CELL_BLOCK_MAX
#define GCDT_CELL_BLOCK_MAX 2048
#endif
constexpr int CELL_RANK_MAX = GCDT_CELL_RANK_MAX;
constexpr int CELL_BLOCK_MAX = GCDT_CELL_BLOCK_MAX;
static_assert(CELL_RANK_MAX <= CELL_BLOCK_MAX, "bucket caps");

// Points: the key is the input index; a position holds (index, cell, x, y),
// so placing a point writes its coordinates, its index and the inverse map
// without gathering the point again.
struct CellKeyPoint {
    struct Key {
        int a;
    };
    using Slot = int4;
    const int2* pts;  // input points (for the paths that place by element)
    int* perm;        // sorted position -> input index
    int2* out;        // sorted points
    int* inv;         // input index -> sorted position
    __device__ Key of(int e) const { return Key{e}; }
    __device__ static Key max_key() { return Key{INT_MAX}; }
    __device__ static bool less(const Key& x, const Key& y) { return x.a < y.a; }
    __device__ void put(long long pos, const Slot& e) const
    {
        perm[pos] = e.x;
        out[pos] = make_int2(e.z, e.w);
        inv[e.x] = (int)pos;
    }
    __device__ void put_elem(long long pos, int el) const
    {
        perm[pos] = el;
        out[pos] = pts[el];
        inv[el] = (int)pos;
    }
};

// Leaves: the key is the triangle's vertex ids, smallest first; a position
// holds (triangle, cell) and placing writes the triangle id and its
// destination.
struct CellKeyTri {
    const int3* tri;
    int* src;   // sorted position -> triangle
    int* dest;  // triangle -> sorted position
    using Key = int3;
    using Slot = int2;
    __device__ void put(long long pos, const Slot& e) const { put_elem(pos, e.x); }
    __device__ void put_elem(long long pos, int el) const
    {
        src[pos] = el;
        dest[el] = (int)pos;
    }
    __device__ Key of(int e) const
    {
        const int3 t = tri[e];
        const int a = min(t.x, min(t.y, t.z)), c = max(t.x, max(t.y, t.z));
        return make_int3(a, t.x + t.y + t.z - a - c, c);
    }
    __device__ static Key max_key() { return make_int3(INT_MAX, INT_MAX, INT_MAX); }
    __device__ static bool less(const Key& x, const Key& y)
    {
        return x.x < y.x || (x.x == y.x && (x.y < y.y || (x.y == y.y && x.z < y.z)));
    }
};

// eb[s] = (element, bucket, ...) at position s; a bucket's range is
// [end[b-1], end[b]) (end: inclusive scan).  Places (KF::put) each element of a
// bucket of up to CELL_RANK_MAX at its rank in the bucket's key order and
// every other element where it is; a bucket's first lane lists a bucket of
// CELL_RANK_MAX + 1 .. CELL_BLOCK_MAX for the workgroup pass.  The workgroup
// first stages the keys of every position its ranked buckets cover (its own
// kBlock positions and the parts of its edge buckets beyond them, at most
// CELL_RANK_MAX - 1 on either side), so the rank loops read only LDS.
template <class KF>
__global__ __launch_bounds__(kBlock) void k_cell_rank(KF kf, const int* end, const typename KF::Slot* eb, long long n,
                                                     int* big, int* nbig)
{
    using Key = typename KF::Key;
    using Slot = typename KF::Slot;
    __shared__ Key sk[kBlock + 2 * (CELL_RANK_MAX - 1)];
    __shared__ int s_lo, s_hi;
    for (long long s0 = (long long)blockIdx.x * kBlock; s0 < n; s0 += (long long)gridDim.x * kBlock) {
        const long long s = s0 + threadIdx.x;
        Slot e{};
        long long a = 0, k = 0;
        if (s < n) {
            e = eb[s];
            a = e.y ? end[e.y - 1] : 0;
            k = end[e.y] - a;
        }
        const bool rank = s < n && k > 1 && k <= CELL_RANK_MAX;
        if (threadIdx.x == 0) {
            s_lo = kBlock;
            s_hi = 0;
        }
        __syncthreads();
        // window [lo, hi) relative to s0: lo >= -(CELL_RANK_MAX - 1),
        // hi <= kBlock + CELL_RANK_MAX - 1
        {
            // one LDS atomic per wave: the wave's own bounds first
            int wl = rank ? (int)(a - s0) : kBlock, wh = rank ? (int)(a + k - s0) : 0;
            for (int d = 32; d > 0; d >>= 1) {
                wl = min(wl, __shfl_down(wl, d, 64));
                wh = max(wh, __shfl_down(wh, d, 64));
            }
            if (__lane_id() == 0 && wl < wh) {
                atomicMin(&s_lo, wl);
                atomicMax(&s_hi, wh);
            }
        }
        __syncthreads();
        const int lo = s_lo, hi = s_hi;
        // a lane stages its own key when its cell is ranked (the keys of
        // unranked positions are never read); the parts of the edge cells
        // beyond the workgroup's positions are staged in turn
        if (rank) sk[(int)threadIdx.x - lo] = kf.of(e.x);
        for (int t = lo + (int)threadIdx.x; t < 0; t += kBlock) sk[t - lo] = kf.of(eb[s0 + t].x);
        for (int t = kBlock + (int)threadIdx.x; t < hi; t += kBlock) sk[t - lo] = kf.of(eb[s0 + t].x);
        __syncthreads();
        bool list = false;
        if (s < n) {
            long long pos = s;
            if (rank) {
                const Key mine = sk[s - s0 - lo];
                const Key* w = sk + (a - s0 - lo);
                int r = 0;
                for (int q = 0; q < (int)k; q++) r += KF::less(w[q], mine) ? 1 : 0;
                pos = a + r;
            }
            kf.put(pos, e);
            list = s == a && k > CELL_RANK_MAX && k <= CELL_BLOCK_MAX;
        }
        const int p = wave_alloc(list ? 1 : 0, nbig);
        if (list) big[p] = e.y;
        __syncthreads();
    }
}

// one workgroup per listed bucket: a bitonic sort in LDS, padded to a power of
// two with largest keys, of the bucket's elements into its range of out[]
template <class KF>
__global__ __launch_bounds__(kBlock) void k_cell_sort_block(KF kf, const int* end, const int* big, const int* nbig,
                                                              const typename KF::Slot* eb)
{
    using Key = typename KF::Key;
    __shared__ int se[CELL_BLOCK_MAX];
    __shared__ Key sk[CELL_BLOCK_MAX];
    const int nb = *nbig;
    for (int q = blockIdx.x; q < nb; q += gridDim.x) {
        const int b = big[q];
        const int a = b ? end[b - 1] : 0, k = end[b] - a;
        int np2 = 1;
        while (np2 < k) np2 <<= 1;
        for (int i = threadIdx.x; i < np2; i += kBlock) {
            const int el = i < k ? eb[a + i].x : -1;
            se[i] = el;
            sk[i] = i < k ? kf.of(el) : KF::max_key();
        }
        for (int size = 2; size <= np2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                __syncthreads();
                for (int i = threadIdx.x; i < (np2 >> 1); i += kBlock) {
                    const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    if (KF::less(sk[hi], sk[lo]) == up) {
                        const Key tk = sk[lo];
                        sk[lo] = sk[hi];
                        sk[hi] = tk;
                        const int te = se[lo];
                        se[lo] = se[hi];
                        se[hi] = te;
                    }
                }
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < k; i += kBlock) kf.put_elem(a + i, se[i]);
        __syncthreads();
    }
}

// Buckets of more than CELL_BLOCK_MAX elements (a dense cluster: the count
// passes flag them, so the common solve never takes this path) are sorted
// together as (bucket, key, element) items: the items of every such bucket,
// compacted in position order, are sorted in chunks of CELL_BLOCK_MAX in LDS
// and then merged pairwise, each item placing itself in the merged run by a
// binary search of the other run (keys are unique); item j of the sorted
// array goes back to the j-th flagged position.
template <class KF>
struct CellItem {
    typename KF::Key key;
    int bucket;
    int elem;
};

template <class KF>
__device__ inline bool item_less(const CellItem<KF>& x, const CellItem<KF>& y)
{
    return x.bucket < y.bucket || (x.bucket == y.bucket && KF::less(x.key, y.key));
}

template <class SLOT>
__global__ __launch_bounds__(kBlock) void k_cell_huge_flags(const int* end, const SLOT* eb, long long n, int* flag)
{
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (long long)gridDim.x * blockDim.x) {
        const int b = eb[s].y;
        flag[s] = end[b] - (b ? end[b - 1] : 0) > CELL_BLOCK_MAX ? 1 : 0;
    }
}

template <class KF>
__global__ __launch_bounds__(kBlock) void k_cell_huge_gather(KF kf, const typename KF::Slot* eb, long long n,
                                                             const int* flag,
                                                             const int* off, CellItem<KF>* items, int* pos)
{
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (long long)gridDim.x * blockDim.x) {
        if (!flag[s]) continue;
        const typename KF::Slot e = eb[s];
        const int j = off[s];
        items[j] = CellItem<KF>{kf.of(e.x), e.y, e.x};
        pos[j] = (int)s;
    }
}

// one workgroup per chunk of CELL_BLOCK_MAX items: bitonic sort in LDS, the
// last chunk padded with items above every real one
template <class KF>
__global__ __launch_bounds__(kBlock) void k_cell_chunk_sort(CellItem<KF>* items, long long m)
{
    using Item = CellItem<KF>;
    __shared__ Item sv[CELL_BLOCK_MAX];
    const long long nchunks = (m + CELL_BLOCK_MAX - 1) / CELL_BLOCK_MAX;
    for (long long q = blockIdx.x; q < nchunks; q += gridDim.x) {
        const long long a = q * CELL_BLOCK_MAX;
        const int k = (int)min((long long)CELL_BLOCK_MAX, m - a);
        int np2 = 1;
        while (np2 < k) np2 <<= 1;
        for (int i = threadIdx.x; i < np2; i += kBlock)
            sv[i] = i < k ? items[a + i] : Item{KF::max_key(), INT_MAX, -1};
        for (int size = 2; size <= np2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                __syncthreads();
                for (int i = threadIdx.x; i < (np2 >> 1); i += kBlock) {
                    const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    if (item_less<KF>(sv[hi], sv[lo]) == up) {
                        const Item t = sv[lo];
                        sv[lo] = sv[hi];
                        sv[hi] = t;
                    }
                }
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < k; i += kBlock) items[a + i] = sv[i];
        __syncthreads();
    }
}

// runs of w sorted items, merged in pairs: an item at index i of its run goes
// to the pair's start + i + (items of the other run below it)
template <class KF>
__global__ __launch_bounds__(kBlock) void k_cell_merge(const CellItem<KF>* in, CellItem<KF>* out, long long m,
                                                       long long w)
{
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (long long)gridDim.x * blockDim.x) {
        const long long run = j / w, base = (run & ~1LL) * w;
        const long long o0 = (run ^ 1) * w, o1 = min(o0 + w, m);
        const CellItem<KF> x = in[j];
        long long lo = o0, hi = o1 > o0 ? o1 : o0;
        while (lo < hi) {
            const long long mid = (lo + hi) >> 1;
            if (item_less<KF>(in[mid], x)) lo = mid + 1;
            else hi = mid;
        }
        out[base + (j - run * w) + (lo - o0)] = x;
    }
}

template <class KF>
__global__ __launch_bounds__(kBlock) void k_cell_huge_put(KF kf, const CellItem<KF>* items, const int* pos, long long m)
{
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (long long)gridDim.x * blockDim.x)
        kf.put_elem(pos[j], items[j].elem);
}

template <class KF>
static void cell_reorder_huge(Ctx& c, const KF& kf, const int* end, const typename KF::Slot* eb, long long n)
{
    DBuf<int> flag((size_t)n), off((size_t)n + 1);  // the scan writes its total at off[n]
    k_cell_huge_flags<typename KF::Slot><<<grid_for(n), kBlock, 0, c.st>>>(end, eb, n, flag.p);
    GCDT_LAUNCH_CHECK();
    scan_excl_total(c, flag.p, off.p, n);
    const long long m = c.get(C_TOTAL);
    if (m <= 0) return;
    DBuf<CellItem<KF>> a((size_t)m), b((size_t)m);
    DBuf<int> pos((size_t)m);
    k_cell_huge_gather<KF><<<grid_for(n), kBlock, 0, c.st>>>(kf, eb, n, flag.p, off.p, a.p, pos.p);
    GCDT_LAUNCH_CHECK();
    flag.release();
    off.release();
    k_cell_chunk_sort<KF><<<1024, kBlock, 0, c.st>>>(a.p, m);
    GCDT_LAUNCH_CHECK();
    CellItem<KF>*x = a.p, *y = b.p;
    for (long long w = CELL_BLOCK_MAX; w < m; w <<= 1) {
        k_cell_merge<KF><<<grid_for(m), kBlock, 0, c.st>>>(x, y, m, w);
        GCDT_LAUNCH_CHECK();
        std::swap(x, y);
    }
    k_cell_huge_put<KF><<<grid_for(m), kBlock, 0, c.st>>>(kf, x, pos.p, m);
    GCDT_LAUNCH_CHECK();
}

// eb[0..n): (element, bucket, ...), grouped by bucket.  Each element is
// placed (KF::put) at its position in the reordered order.  huge: some bucket
// holds more than CELL_BLOCK_MAX elements.
template <class KF>
static void cell_reorder(Ctx& c, const KF& kf, const int* end, const typename KF::Slot* eb, long long n, bool huge)
{
    if (n <= 0) return;
    // a list entry per bucket above the rank cap: at most n / (cap + 1)
    DBuf<int> big((size_t)(n / (CELL_RANK_MAX + 1) + 1));
    c.set(C_NQ, 0);
    k_cell_rank<KF><<<grid_for(n), kBlock, 0, c.st>>>(kf, end, eb, n, big.p, c.dcnt.p + C_NQ);
    GCDT_LAUNCH_CHECK();
    k_cell_sort_block<KF><<<1024, kBlock, 0, c.st>>>(kf, end, big.p, c.dcnt.p + C_NQ, eb);
    GCDT_LAUNCH_CHECK();
    if (huge) cell_reorder_huge(c, kf, end, eb, n);
}

// Counting sort of points by 20-bit Morton cell: the count pass keeps the
// value its atomicAdd returns (the point's arrival rank in its cell) in
// inv[], the scatter writes (index, cell, x, y) at the cell's start plus that
// rank, and the per-cell reorder places each point by input index, writing
// the sorted points, perm and the inverse map.
__global__ __launch_bounds__(kBlock) void k_morton_count(const int2* pts, int n, int* cnt, int* rank, int* huge)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int r = atomicAdd(&cnt[morton_cell(pts[i])], 1);
        rank[i] = r;
        if (r == CELL_BLOCK_MAX) atomicAdd(huge, 1);
    }
}

// cell_end holds the inclusive scan of the counts: cell c starts at
// cell_end[c - 1] (0 for c = 0).
__global__ __launch_bounds__(kBlock) void k_morton_scatter(const int2* pts, int n, const int* cell_end,
                                                           const int* rank, int4* eb)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int2 p = pts[i];
        const int cell = (int)morton_cell(p);
        const int pos = (cell ? cell_end[cell - 1] : 0) + rank[i];
        eb[pos] = make_int4((int)i, cell, p.x, p.y);
    }
}

void morton_order(Ctx& c, const int2* pts, int n, int2* out, int* perm, int* inv, int* cell_end)
{
    fill_i32(c, cell_end, 0, BITMAP_CELLS + 1);
    if (n <= 0) return;
    c.set(C_HUGE, 0);
    k_morton_count<<<grid_for(n), kBlock, 0, c.st>>>(pts, n, cell_end, inv, c.dcnt.p + C_HUGE);
    GCDT_LAUNCH_CHECK();
    scan_incl(c, cell_end, cell_end, BITMAP_CELLS);
    DBuf<int4> eb((size_t)n);
    k_morton_scatter<<<grid_for(n), kBlock, 0, c.st>>>(pts, n, cell_end, inv, eb.p);
    GCDT_LAUNCH_CHECK();
    // inv held the arrival ranks; placing rewrites it as the inverse map
    cell_reorder(c, CellKeyPoint{pts, perm, out, inv}, cell_end, eb.p, n, c.get(C_HUGE) > 0);
}

// Sample of Morton-ordered points: the 256x256 sample cell with Morton index
// m covers the 16 consecutive 20-bit cells [16m, 16m+16), i.e. the sorted
// range [end[16m-1], end[16m+15]).  Its sample is its extreme point towards
// the nearer edges of the square: the smallest (x, y) in lexicographic order,
// with x taken as -x in the right half of the square and y as -y in the top
// half.  Any point of the cell serves location; the extreme one does not
// depend on the order of the sorted range, and on collinear rows and columns
// -- a lattice, or points clamped to the square's edge -- it puts the samples
// on the outermost lines, so the points between two samples split the edge
// joining them instead of fanning out from a sample off the line (a fan that
// Lawson flips unwind one row per round).
__device__ inline int sample_range_start(const int* cell_end, int m)
{
    return m ? cell_end[16 * m - 1] : 0;
}

__global__ __launch_bounds__(kBlock) void k_sample_morton_flags(const int* cell_end, int* flags)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < SAMPLE_CELLS; i += gridDim.x * blockDim.x)
        flags[i] = cell_end[16 * i + 15] > sample_range_start(cell_end, i);
}

// one wave per sample cell: its extreme point over the cell's sorted range
__global__ __launch_bounds__(kBlock) void k_sample_morton_emit(const int2* pts, const int* cell_end, const int* offs,
                                                               int* ids)
{
    const int lane = __lane_id();
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int m = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; m < SAMPLE_CELLS; m += nwaves) {
        const int a = sample_range_start(cell_end, m), b = cell_end[16 * m + 15];
        if (b <= a) continue;  // wave-uniform
        // the cell's column and row (even and odd bits of its Morton index)
        unsigned cx = 0, cy = 0;
        for (int bit = 0; bit < 8; bit++) {
            cx |= (((unsigned)m >> (2 * bit)) & 1u) << bit;
            cy |= (((unsigned)m >> (2 * bit + 1)) & 1u) << bit;
        }
        const unsigned fx = cx >= 128 ? ~0u : 0u, fy = cy >= 128 ? ~0u : 0u;
        unsigned long long best = ~0ull;
        int bi = a;
        for (int i = a + lane; i < b; i += 64) {
            const int2 p = pts[i];
            const unsigned long long k = ((unsigned long long)((unsigned)p.x ^ fx) << 32) | ((unsigned)p.y ^ fy);
            if (k < best) {
                best = k;
                bi = i;
            }
        }
        unsigned long long w = best;
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl(w, lane ^ off, 64);
            w = o < w ? o : w;
        }
        // points are distinct: exactly one lane holds the minimum
        const unsigned long long owner = __ballot(best == w);
        const int src = __ffsll((long long)owner) - 1;
        const int id = __shfl(bi, src, 64);
        if (lane == 0) ids[offs[m]] = id;
    }
}

int sample_points_morton(Ctx& c, const int2* pts, const int* cell_end, int* sample_ids)
{
    DBuf<int> offs(SAMPLE_CELLS + 1);
    k_sample_morton_flags<<<grid_for(SAMPLE_CELLS), kBlock, 0, c.st>>>(cell_end, offs.p);
    GCDT_LAUNCH_CHECK();
    scan_excl(c, offs.p, offs.p, SAMPLE_CELLS);
    // one wave per sample cell
    k_sample_morton_emit<<<grid_for(64LL * SAMPLE_CELLS), kBlock, 0, c.st>>>(pts, cell_end, offs.p, sample_ids);
    GCDT_LAUNCH_CHECK();
    return c.get(C_TOTAL);
}

__global__ __launch_bounds__(kBlock) void k_relabel_pairs(const int2* in, int C, const int* inv, int n, int2* out)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < C; i += (long long)gridDim.x * blockDim.x) {
        int2 e = in[i];
        // ids outside [0,n) are kept and rejected later by the constraint checks
        e.x = (e.x >= 0 && e.x < n) ? inv[e.x] : e.x;
        e.y = (e.y >= 0 && e.y < n) ? inv[e.y] : e.y;
        out[i] = e;
    }
}

__global__ __launch_bounds__(kBlock) void k_unmap_tris(int3* tri, int T, const int* perm, int n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        int3 t = tri[i];
        t.x = t.x < n ? perm[t.x] : t.x;
        t.y = t.y < n ? perm[t.y] : t.y;
        t.z = t.z < n ? perm[t.z] : t.z;
        tri[i] = t;
    }
}

void relabel_pairs(Ctx& c, const int2* in, int C, const int* inv, int n, int2* out)
{
    if (C <= 0) return;
    k_relabel_pairs<<<grid_for(C), kBlock, 0, c.st>>>(in, C, inv, n, out);
    GCDT_LAUNCH_CHECK();
}

void unmap_tris(Ctx& c, int3* tri, int T, const int* perm, int n)
{
    if (T <= 0 || n <= 0) return;
    k_unmap_tris<<<grid_for(T), kBlock, 0, c.st>>>(tri, T, perm, n);
    GCDT_LAUNCH_CHECK();
}

// ================================================================== location

// initTriangle, rawMesh.cu:37-57: side of the anti-diagonal (R,0)-(0,R)
__global__ __launch_bounds__(kBlock) void k_locate_two(const int2* pts, int n, int* loc, int* pend)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        int o = orient2d(make_int2(GRID_R, 0), make_int2(0, GRID_R), pts[i]);
        loc[i] = o > 0 ? 3 : (o < 0 ? 4 + 3 : 0);  // tri 0 inside / tri 1 inside / edge opposite vertex 0 of tri 0
        pend[i] = (int)i;
    }
}

__global__ __launch_bounds__(kBlock) void k_locate_two_ids(const int2* pts, const int* ids, int n, int* loc, int* pend)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        int p = ids[i];
        int o = orient2d(make_int2(GRID_R, 0), make_int2(0, GRID_R), pts[p]);
        loc[p] = o > 0 ? 3 : (o < 0 ? 4 + 3 : 0);
        pend[i] = p;
    }
}

int locate_two_ids(Ctx& c, const int2* pts, const int* ids, int n, int* loc, int* pend)
{
    if (n > 0) {
        k_locate_two_ids<<<grid_for(n), kBlock, 0, c.st>>>(pts, ids, n, loc, pend);
        GCDT_LAUNCH_CHECK();
    }
    return n;
}

int locate_two(Ctx& c, const int2* pts, int n, int* loc, int* pend)
{
    if (n > 0) {
        k_locate_two<<<grid_for(n), kBlock, 0, c.st>>>(pts, n, loc, pend);
        GCDT_LAUNCH_CHECK();
    }
    return n;
}

__global__ __launch_bounds__(kBlock) void k_mark_vertices(const int3* tri, int T, int n, int* isv)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        int3 t = tri[i];
        if (t.x < n) isv[t.x] = 1;
        if (t.y < n) isv[t.y] = 1;
        if (t.z < n) isv[t.z] = 1;
    }
}

// initTriangleBoost, rawMesh.cu:59-152: scan the triangles listed for the
// point's bitmap cell; the first whose closed region holds the point wins.
template <class G>
__global__ __launch_bounds__(kBlock) void k_locate_bitmap(G g, const int2* pts, int n, const int* isv, const int3* tri,
                                                          const int* bm_off, const int* bm_tri, int* loc, int* pend,
                                                          int* npend, int* err)
{
    __shared__ int s_buf[STAGE_CAP];
    __shared__ int s_base;
    Staged<STAGE_CAP> sp(s_buf, &s_base, pend, npend);
    TILE_LOOP(base, n, kBlock)
    {
        long long i = base + threadIdx.x;
        bool want = false;
        if (i < n) {
            int p = (int)i;
            if (isv[p]) {
                loc[p] = -1;
            } else {
                int2 P = pts[p];
                long long cell = pt_cell(g, p, P);
                int r = -1, t = -1;
                for (int j = bm_off[cell], e = bm_off[cell + 1]; j < e; j++) {
                    t = bm_tri[j];
                    int3 tv = tri[t];
                    r = classify(pts[tv.x], pts[tv.y], pts[tv.z], P);
                    if (r >= 0 || r == -2) break;
                }
                if (r >= 0) {
                    loc[p] = t * 4 + r;
                    want = true;
                } else {
                    loc[p] = -1;
                    atomicOr(err, r == -2 ? MERR_DUPLICATE : MERR_LOCATE);
                }
            }
        }
        int slot = sp.reserve(want ? 1 : 0);
        if (want) s_buf[slot] = (int)i;
    }
    sp.flush();
}

template <class G>
static int locate_bitmap_g(Ctx& c, const G& g, const int2* pts, int n, const int3* tri, int T, const int* bm_off,
                           const int* bm_tri, int* loc, int* isv, int* pend)
{
    k_mark_vertices<<<grid_for(T), kBlock, 0, c.st>>>(tri, T, n, isv);
    GCDT_LAUNCH_CHECK();
    c.set(C_NPEND, 0);
    if (n > 0) {
        ProfScope ps("locate_bitmap", c.st, n);
        k_locate_bitmap<G><<<grid_for(n), kBlock, 0, c.st>>>(g, pts, n, isv, tri, bm_off, bm_tri, loc, pend,
                                                             c.dcnt.p + C_NPEND, c.dcnt.p + C_ERR);
        GCDT_LAUNCH_CHECK();
    }
    c.check_err();
    return c.get(C_NPEND);
}

int locate_bitmap(Ctx& c, const int2* pts, int n, const int3* tri, int T, const int* bm_off, const int* bm_tri,
                  int* loc, int* isv, int* pend)
{
    return locate_bitmap_g(c, GridOne{}, pts, n, tri, T, bm_off, bm_tri, loc, isv, pend);
}

int locate_bitmap_batch(Ctx& c, const int2* pts, const int* vinst, long long NP, const int3* tri, int T,
                        const int* bm_off, const int* bm_tri, int* loc, int* isv, int* pend)
{
    return locate_bitmap_g(c, GridInst{vinst, NP}, pts, (int)NP, tri, T, bm_off, bm_tri, loc, isv, pend);
}

// ================================================================== insertion rounds

// Round prologue on the device: the previous round's output list becomes the
// input (count C_NPEND2 -> C_NPEND) and the slot base is recorded.
// With snap set, it first copies the previous round's counters to the host
// ring (one launch instead of a begin kernel plus a copy).  C_NUM threads.
__global__ void k_round_begin(int* cnt, int move, unsigned long long* snap, unsigned seq)
{
    if (snap) snap_store(&snap[threadIdx.x], seq, cnt[threadIdx.x]);
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (move) cnt[C_NPEND] = cnt[C_NPEND2];
    cnt[C_NPEND2] = 0;
    cnt[C_TBASE] = cnt[C_NTRI];
    cnt[C_NW] = cnt[C_NW2];
    cnt[C_NW2] = 0;
}

// The claimed-triangle list of a round: the claim whose atomicMax finds a key
// of an older epoch is the triangle's first this round and appends it, so
// every claimed triangle is listed once and the split step walks these
// triangles instead of every pending point.
__device__ inline bool claim_first(unsigned long long old, unsigned epoch)
{
    return (unsigned)(old >> 44) != (epoch & 0xfffffu);
}

// Wave-combined claims: consecutive lanes that target the same triangle form
// a run (the pending list stays in near-Morton order, so early rounds have
// long runs); a segmented max over each run leaves the run's highest key in
// its first lane, which issues the run's only atomic.  *f1 / *f2 report
// whether this lane's claims on t / t2 came first this round.  All lanes must
// be active.
__device__ inline void claim_combined_w(unsigned long long* claim, int t, int t2, unsigned long long key,
                                        unsigned epoch, bool* f1, bool* f2)
{
    const int lane = __lane_id();
    const int prev = __shfl_up(t, 1, 64);
    const bool head = lane == 0 || prev != t;
    const unsigned long long heads = __ballot(head);
    const unsigned long long after = heads & ~((2ull << lane) - 1);
    const int end = after ? __ffsll((long long)after) - 2 : 63;  // last lane of this lane's run
    unsigned long long v = t >= 0 ? key : 0ull;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned long long o = __shfl_down(v, d, 64);
        if (lane + d <= end) v = o > v ? o : v;
    }
    *f1 = head && t >= 0 && claim_first(atomicMax(&claim[t], v), epoch);
    *f2 = t2 >= 0 && claim_first(atomicMax(&claim[t2], key), epoch);
}

// Claims of `epoch` for up to IT points per thread (p < 0: none), appending
// first-claimed triangles to wl.  Block-uniform: every thread of the block
// calls it.
template <int IT>
__device__ inline void claim_items(const Ins& s, const int (&pp)[IT], unsigned epoch, Staged<STAGE_CAP>& sw)
{
    int wt[2 * IT];
    int nq = 0;
#pragma unroll
    for (int j = 0; j < IT; j++) {
        int t = -1, t2 = -1;
        unsigned long long key = 0;
        if (pp[j] >= 0) t = ins_key_round(s, pp[j], epoch, &key, &t2);
        bool f1, f2;
        claim_combined_w(s.claim, t, t2, key, epoch, &f1, &f2);
        wt[2 * j] = f1 ? t : -1;
        wt[2 * j + 1] = f2 ? t2 : -1;
        nq += (wt[2 * j] >= 0) + (wt[2 * j + 1] >= 0);
    }
    int b = sw.reserve(nq);
#pragma unroll
    for (int j = 0; j < 2 * IT; j++)
        if (wt[j] >= 0) sw.buf[b++] = wt[j];
}

// First claims of a run of rounds (later rounds claim inside k_ins_reloc).
__global__ __launch_bounds__(kBlock) void k_ins_claim_w(Ins s, const int* list, const int* cnt, unsigned epoch, int* wl,
                                                        int* nw, int wcap)
{
    __shared__ int s_buf[STAGE_CAP];
    __shared__ int s_base;
    Staged<STAGE_CAP> sw(s_buf, &s_base, wl, nw, wcap, s.err, MERR_CAPACITY);
    int n = *cnt;
    TILE_LOOP(base, n, kBlock)
    {
        long long i = base + threadIdx.x;
        int pp[1] = {i < n ? list[i] : -1};
        claim_items<1>(s, pp, epoch, sw);
    }
    sw.flush();
}

// Items per thread of the block-appending list kernels: 1, 2 or 4 (tunable;
// more items amortise the block-wide append, fewer keep more waves in flight).
template <class F>
void with_items(int v, F&& f)
{
    if (v == 1) f(std::integral_constant<int, 1>{});
    else if (v == 2) f(std::integral_constant<int, 2>{});
    else f(std::integral_constant<int, 4>{});
}

// Splits for the round's claimed triangles: a triangle's claim word names its
// winning point; an on-edge winner splits from its own triangle (loc), not
// from the second one it claimed.
template <int INS_ITEMS>
__global__ __launch_bounds__(kBlock) void k_ins_split(Ins s, const int* list, const int* cnt, int lcap, unsigned epoch,
                                                      int* ntri)
{
    int n = min(*cnt, lcap);  // an overfull list has flagged MERR_CAPACITY
    TILE_LOOP(base, n, kBlock * INS_ITEMS)
    {
        int pp[INS_ITEMS], need[INS_ITEMS];
        int tot = 0;
#pragma unroll
        for (int j = 0; j < INS_ITEMS; j++) {
            long long i = base + (long long)j * kBlock + threadIdx.x;
            pp[j] = -1;
            need[j] = 0;
            if (i < n) {
                int t = list[i];
                int p = (int)(unsigned)s.claim[t];
                int l = s.loc[p];
                if (l >= 0 && (l >> 2) == t) {
                    pp[j] = p;
                    need[j] = ins_need<false>(s, p, epoch);
                }
            }
            tot += need[j];
        }
        int b = block_alloc<kBlock>(tot, ntri);
#pragma unroll
        for (int j = 0; j < INS_ITEMS; j++) {
            if (need[j]) {
                if (b + need[j] > s.cap) atomicOr(s.err, MERR_CAPACITY);
                else ins_split(s, pp[j], b);
                b += need[j];
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_ins_fix(Ins s, const int* cnt)
{
    int b = cnt[C_TBASE], e = min(cnt[C_NTRI], s.cap);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x + b; i < e; i += (long long)gridDim.x * blockDim.x)
        ins_fix(s, (int)i);
}

// Relocation of the pending points after the splits, fused with the next
// round's claims (epoch next_ep) and its claimed-triangle list.
template <int INS_ITEMS>
__global__ __launch_bounds__(kBlock) void k_ins_reloc(Ins s, const int* list, const int* cnt, int* list2, int* cnt2,
                                                      unsigned next_ep, int* wl, int* nw, int wcap)
{
    __shared__ int s_buf[STAGE_CAP], s_buf2[STAGE_CAP];
    __shared__ int s_base, s_base2;
    Staged<STAGE_CAP> sl(s_buf, &s_base, list2, cnt2), sw(s_buf2, &s_base2, wl, nw, wcap, s.err, MERR_CAPACITY);
    int n = *cnt;
    TILE_LOOP(base, n, kBlock * INS_ITEMS)
    {
        int keep[INS_ITEMS];
        int tot = 0;
#pragma unroll
        for (int j = 0; j < INS_ITEMS; j++) {
            long long i = base + (long long)j * kBlock + threadIdx.x;
            keep[j] = -1;
            if (i < n) {
                int p = list[i];
                if (ins_relocate(s, p)) keep[j] = p;
            }
            tot += keep[j] >= 0;
        }
        int b = sl.reserve(tot);
#pragma unroll
        for (int j = 0; j < INS_ITEMS; j++)
            if (keep[j] >= 0) s_buf[b++] = keep[j];
        claim_items<INS_ITEMS>(s, keep, next_ep, sw);
    }
    sl.flush();
    sw.flush();
}

// ---- persistent single-workgroup tail: all remaining insertion rounds in one
// launch (block barriers between the phases, LDS counters for the lists)
constexpr int TAIL_BS = 1024;

__global__ __launch_bounds__(TAIL_BS) void k_ins_tail(Ins s, int* P, int* P2, int* cnt, int cp, unsigned epoch0,
                                                      int max_rounds)
{
    __shared__ int s_np, s_nt, s_tb;
    int n = cnt[cp];
    unsigned ep = epoch0;
    int r = 0, err = 0;
    if (threadIdx.x == 0) s_nt = __hip_atomic_load(&cnt[C_NTRI], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    while (n > 0 && r < max_rounds) {
        for (int i = threadIdx.x; i < n; i += TAIL_BS) ins_claim(s, P[i], ep);
        if (threadIdx.x == 0) {
            s_np = 0;
            s_tb = s_nt;
        }
        __syncthreads();
        for (int i0 = 0; i0 < n; i0 += TAIL_BS) {
            const int i = i0 + (int)threadIdx.x;
            const int p = i < n ? P[i] : -1;
            const int need = p >= 0 ? ins_need(s, p, ep) : 0;
            const int b = wave_alloc(need, &s_nt);
            if (need) {
                if (b + need > s.cap) atomicOr(s.err, MERR_CAPACITY);
                else ins_split(s, p, b);
            }
        }
        __syncthreads();
        int tb = s_tb, te = min(s_nt, s.cap);
        for (int c = tb + (int)threadIdx.x; c < te; c += TAIL_BS) ins_fix(s, c);
        __syncthreads();
        for (int i0 = 0; i0 < n; i0 += TAIL_BS) {
            const int i = i0 + (int)threadIdx.x;
            const int p = i < n ? P[i] : -1;
            const bool keep = p >= 0 && ins_relocate(s, p);
            const int b = wave_alloc(keep ? 1 : 0, &s_np);
            if (keep) P2[b] = p;
        }
        __syncthreads();
        int n2 = s_np;
        if (n2 >= n && te == tb) {
            err = 1;
            break;
        }
        n = n2;
        int* tmp = P;
        P = P2;
        P2 = tmp;
        ep++;
        r++;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cnt[C_NTRI] = s_nt;
        cnt[C_TAIL_N] = n;
        cnt[C_TAIL_R] = r;
        cnt[C_TAIL_ERR] = err;
        cnt[C_TAIL_PAR] = r & 1;
    }
}

int insert_rounds(Ctx& c, const Ins& s_in, DBuf<int>& pend, DBuf<int>& pend2, int npend, int ntri, int* rounds)
{
    Ins s = s_in;
    s.pivot = c.pivot;
    c.set(C_NPEND, npend);
    c.set(C_NTRI, ntri);
    int* P = pend.p;   // input list of the next round to launch
    int* P2 = pend2.p;
    int r = 0;
    int est = npend;  // estimate of the next round's input size (exact for the first)
    bool done = npend == 0;
    // Rounds are launched one ahead of the host's knowledge of their input
    // size: each round's counters are snapshot asynchronously and read after
    // the next round is queued, so the GPU does not wait for the host.
    auto check_round = [&](const int* h) {
        if (h[C_ERR]) c.check_err();
        if (h[C_NPEND] > 0 && h[C_NPEND2] >= h[C_NPEND] && h[C_NTRI] == h[C_TBASE])
            throw Error(GCDT_ERR_INTERNAL, "insertion round without progress");
        // exact list lengths of the round, for per-item bytes of the timed kernels
        prof_add_items("ins_split", h[C_NW]);
        prof_add_items("ins_reloc", h[C_NPEND]);
    };
    // Claimed-triangle lists: every pending point claims at most two
    // triangles, and each claimed triangle is listed once per round.
    DBuf<int> wa, wb;
    int* W = nullptr;   // triangles claimed for the round being launched
    int* W2 = nullptr;  // triangles claimed for the round after it
    long long wcap = std::min<long long>(2LL * npend, s.cap) + 1;
    if (!done && est > c.tail_threshold) {
        wa.ensure((size_t)wcap);
        wb.ensure((size_t)wcap);
        W = wa.p;
        W2 = wb.p;
        // the first round's claims; later rounds claim inside the relocation
        c.set(C_NW2, 0);
        ProfScope ps("ins_claim", c.st, est);
        k_ins_claim_w<<<grid_for(est), kBlock, 0, c.st>>>(s, P, c.dcnt.p + C_NPEND, c.epoch + 1, W2,
                                                          c.dcnt.p + C_NW2, (int)wcap);
        GCDT_LAUNCH_CHECK();
    }
    while (!done && est > c.tail_threshold) {
        if (++r > 100000) throw Error(GCDT_ERR_NOT_CONVERGED, "insertion rounds did not converge");
        unsigned ep = c.next_epoch();
        // begins round r and snapshots round r-1 for the host
        unsigned long long* snap = r > 1 ? c.snap_dst((r - 1) & 1) : nullptr;
        unsigned seq = r > 1 ? c.snap_begin((r - 1) & 1) : 0u;
        k_round_begin<<<1, C_NUM, 0, c.st>>>(c.dcnt.p, r > 1, snap, seq);
        GCDT_LAUNCH_CHECK();
        std::swap(W, W2);
        with_items(c.ins_items, [&](auto it) {
            constexpr int IT = decltype(it)::value;
            ProfScope ps("ins_split", c.st, 0);
            k_ins_split<IT><<<grid_for((est + IT - 1) / IT), kBlock, 0, c.st>>>(s, W, c.dcnt.p + C_NW, (int)wcap,
                                                                                 ep, c.dcnt.p + C_NTRI);
            GCDT_LAUNCH_CHECK();
        });
        {
            ProfScope ps("ins_fix", c.st, 0);
            k_ins_fix<<<grid_for(4LL * est), kBlock, 0, c.st>>>(s, c.dcnt.p);
            GCDT_LAUNCH_CHECK();
        }
        with_items(c.ins_items, [&](auto it) {
            constexpr int IT = decltype(it)::value;
            ProfScope ps("ins_reloc", c.st, 0);
            k_ins_reloc<IT><<<grid_for((est + IT - 1) / IT), kBlock, 0, c.st>>>(
                s, P, c.dcnt.p + C_NPEND, P2, c.dcnt.p + C_NPEND2, ep + 1, W2, c.dcnt.p + C_NW2, (int)wcap);
            GCDT_LAUNCH_CHECK();
        });
        std::swap(P, P2);
        if (r >= 2) {
            const int* h = c.wait_snapshot((r - 1) & 1);  // round r-1 (round r is queued)
            check_round(h);
            if (h[C_NPEND2] == 0) done = true;  // round r had no input
            est = h[C_NPEND2];
        }
    }
    if (r >= 1) {
        c.snapshot(r & 1);
        const int* h = c.wait_snapshot(r & 1);
        check_round(h);
        if (h[C_NPEND2] == 0) done = true;
        est = h[C_NPEND2];
    }
    if (!done) {
        // finish in one persistent workgroup; its input is P with the count moved
        k_round_begin<<<1, C_NUM, 0, c.st>>>(c.dcnt.p, r > 0, nullptr, 0u);
        GCDT_LAUNCH_CHECK();
        unsigned ep0 = c.epoch + 1;
        {
            ProfScope ps("ins_tail", c.st, est);
            k_ins_tail<<<1, TAIL_BS, 0, c.st>>>(s, P, P2, c.dcnt.p, C_NPEND, ep0, 1 << 20);
            GCDT_LAUNCH_CHECK();
        }
        c.read(0, C_NUM);
        if (c.hcnt[C_ERR]) c.check_err();
        if (c.hcnt[C_TAIL_ERR]) throw Error(GCDT_ERR_INTERNAL, "insertion round without progress (tail)");
        r += c.hcnt[C_TAIL_R];
        c.epoch += (unsigned)c.hcnt[C_TAIL_R];
        if (c.hcnt[C_TAIL_N] != 0) throw Error(GCDT_ERR_NOT_CONVERGED, "insertion tail did not converge");
    }
    if (rounds) *rounds = r;
    int nt = c.get(C_NTRI);
    if (nt > s.cap) throw Error(GCDT_ERR_CAPACITY, "triangle capacity exceeded");
    return nt;
}

// ================================================================== flip rounds

// Flip round prologue: the previous round's output list becomes the input.
// With snap set, it first copies the previous round's counters to the host
// ring.  C_NUM threads.
__global__ void k_flip_begin(int* cnt, int move, unsigned long long* snap, unsigned seq)
{
    if (snap) snap_store(&snap[threadIdx.x], seq, cnt[threadIdx.x]);
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (move) cnt[C_NACT] = cnt[C_NACT2];
    cnt[C_NACT2] = 0;
    cnt[C_NFLIP] = 0;
    cnt[C_NPROP] = 0;
}

// Claims of the short rounds of one persistent workgroup, in an LDS hash
// table instead of the records' claim words (device-scope atomics and
// atomic loads there cost a trip to the memory side each; these rounds are
// chains of such trips).  Open addressing on the triangle id.  Tags (-1) and
// keys (0) of a table are cleared before the round that uses it.
constexpr int LCL_CAP = 2048;      // entries per table
constexpr int LCL_MAX_LIST = 512;  // rounds of at most this many items use it (2 claims per item)

struct LdsClaims {
    int* tag;
    unsigned long long* key;
    __device__ void max(const FlipP& f, int t, unsigned long long k) const
    {
        unsigned h = hash32((uint32_t)t) & (LCL_CAP - 1);
        for (int probe = 0; probe < LCL_CAP; probe++) {
            int g = tag[h];
            if (g == -1) g = atomicCAS(&tag[h], -1, t);
            if (g == -1 || g == t) {
                atomicMax(&key[h], k);
                return;
            }
            h = (h + 1) & (LCL_CAP - 1);
        }
        atomicOr(f.err, MERR_CAPACITY);
    }
    __device__ unsigned long long find(int t) const
    {
        unsigned h = hash32((uint32_t)t) & (LCL_CAP - 1);
        for (int probe = 0; probe < LCL_CAP; probe++) {
            int g = tag[h];
            if (g == t) return key[h];
            if (g == -1) return 0ull;
            h = (h + 1) & (LCL_CAP - 1);
        }
        return 0ull;
    }
    template <bool COH>
    __device__ unsigned long long get(const FlipP&, int t, unsigned long long) const
    {
        return find(t);
    }
    template <bool COH>
    __device__ unsigned long long load(const FlipP&, int t) const
    {
        return find(t);
    }
};

// Flip engines: the array-per-field layout (short lists on a mesh whose
// arrays stay in place) and the packed 64-byte records (whole-mesh rounds).
// Both keep the same claims protocol; apply() sets mark = epoch + 1 on the
// two triangles of a flip.
struct EngSoA {
    static constexpr bool kFused = false;
    Flip f;
    __device__ bool test(int t, unsigned ep, bool all, bool single) const { return flip_test(f, t, ep, all, single); }
    __device__ bool fixtest(int t, unsigned ep, bool all, bool single) const { return test(t, ep, all, single); }
    // no LDS claims for this layout (the persistent tail never asks for them)
    __device__ bool fixtest_l(int t, unsigned ep, const LdsClaims&) const { return test(t, ep, false, false); }
    __device__ int apply_l(int t, unsigned ep, int* nn, const LdsClaims&) const { return apply<true>(t, ep, nn); }
    template <bool COH>
    __device__ int apply(int t, unsigned ep, int* nn) const
    {
        int res = flip_apply<COH>(f, t, ep, nn);
        if (res == 2) {
            f.mark[t] = (int)ep + 1;
            f.mark[*nn] = (int)ep + 1;
        }
        return res;
    }
    __device__ void fix(int t, unsigned ep) const { flip_fix(f, t, ep); }
    __device__ void set_mark(int t, int v) const { f.mark[t] = v; }
};

struct EngPacked {
    // the test launch also settles the previous round's flips (flip_fixtest_p)
    static constexpr bool kFused = true;
    FlipP f;
    __device__ bool test(int t, unsigned ep, bool all, bool single) const
    {
        return flip_test_p(f, t, ep, all, single);
    }
    __device__ bool fixtest(int t, unsigned ep, bool all, bool single) const
    {
        return flip_fixtest_p(f, t, ep, all, single);
    }
    __device__ bool fixtest_l(int t, unsigned ep, const LdsClaims& cl) const
    {
        return flip_fixtest_p(f, t, ep, false, false, cl);
    }
    __device__ int apply_l(int t, unsigned ep, int* nn, const LdsClaims& cl) const
    {
        return flip_apply_p<true>(f, t, ep, nn, cl);
    }
    template <bool COH>
    __device__ int apply(int t, unsigned ep, int* nn) const
    {
        return flip_apply_p<COH>(f, t, ep, nn);
    }
    __device__ void fix(int t, unsigned ep) const { flip_fix_p(f, t, ep); }
    __device__ void set_mark(int t, int v) const { f.r[t].meta = rec_with_mark(f.r[t].meta, v); }
};

// Tests the listed triangles; the ones that propose a flip are appended to
// `prop`, the only list the apply step reads.
template <class E, int FT_ITEMS>
__global__ __launch_bounds__(kBlock) void k_flip_test(E e, const int* list, const int* cnt, int lcap, unsigned epoch,
                                                      int single, int* prop, int* nprop, int cap, int* eflag)
{
    __shared__ int s_buf[STAGE_CAP];
    __shared__ int s_base;
    // a triangle proposes at most once per round, so `prop` (cap entries)
    // cannot overflow; a miscount flags MERR_CAPACITY instead of writing past it
    Staged<STAGE_CAP> sp(s_buf, &s_base, prop, nprop, cap, eflag, MERR_CAPACITY);
    int n = min(*cnt, lcap);
    TILE_LOOP(base, n, kBlock * FT_ITEMS)
    {
        int q[FT_ITEMS];
        int nq = 0;
#pragma unroll
        for (int j = 0; j < FT_ITEMS; j++) {
            q[j] = -1;
            long long i = base + (long long)j * kBlock + threadIdx.x;
            if (i < n) {
                int t = list ? list[i] : (int)i;
                if (E::kFused ? e.fixtest(t, epoch, list == nullptr, single != 0)
                              : e.test(t, epoch, list == nullptr, single != 0))
                    q[j] = t;
            }
            nq += q[j] >= 0;
        }
        int b = sp.reserve(nq);
#pragma unroll
        for (int j = 0; j < FT_ITEMS; j++)
            if (q[j] >= 0) s_buf[b++] = q[j];
    }
    sp.flush();
}

template <class E, int FA_ITEMS>
__global__ __launch_bounds__(kBlock) void k_flip_apply(E e, const int* list, const int* cnt, unsigned epoch, int* list2,
                                                       int* cnt2, int* flipped, int* nflip, int cap, int* eflag)
{
    __shared__ int s_buf[STAGE_CAP], s_buf2[STAGE_CAP];
    __shared__ int s_base, s_base2;
    // each triangle enters the next list at most once per round (a lost
    // proposer that keeps its own claim is flipped by no one), so both lists
    // hold at most cap entries; a miscount flags MERR_CAPACITY
    Staged<STAGE_CAP> sq(s_buf, &s_base, list2, cnt2, cap, eflag, MERR_CAPACITY),
        sf(s_buf2, &s_base2, flipped, nflip, cap, eflag, MERR_CAPACITY);
    int n = min(*cnt, cap);
    TILE_LOOP(base, n, kBlock * FA_ITEMS)
    {
        int qa[FA_ITEMS], qb[FA_ITEMS];
        int nq = 0, nf = 0;
#pragma unroll
        for (int j = 0; j < FA_ITEMS; j++) {
            qa[j] = -1;
            qb[j] = -1;
            long long i = base + (long long)j * kBlock + threadIdx.x;
            if (i < n) {
                int t = list[i];
                int nn = -1;
                int res = e.template apply<false>(t, epoch, &nn);
                if (res == 1) e.set_mark(t, (int)epoch + 1);
                if (res == 1 || res == 2) qa[j] = t;
                if (res == 2) qb[j] = nn;
            }
            nq += (qa[j] >= 0) + (qb[j] >= 0);
            nf += (qb[j] >= 0) ? 2 : 0;
        }
        int b = sq.reserve(nq);
        int fb = sf.reserve(nf);
#pragma unroll
        for (int j = 0; j < FA_ITEMS; j++) {
            if (qa[j] >= 0) s_buf[b++] = qa[j];
            if (qb[j] >= 0) {
                s_buf[b++] = qb[j];
                s_buf2[fb++] = qa[j];
                s_buf2[fb++] = qb[j];
            }
        }
    }
    sq.flush();
    sf.flush();
}

template <class E>
__global__ __launch_bounds__(kBlock) void k_flip_fix(E e, const int* list, const int* cnt, int cap, unsigned epoch)
{
    int n = min(*cnt, cap);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        e.fix(list[i], epoch);
}

// Persistent single-workgroup tail of the flip rounds.
template <class E>
__global__ __launch_bounds__(TAIL_BS) void k_flip_tail(E e, int* A, int* B, int* lf, int* cnt, int cap,
                                                       unsigned epoch0, int max_rounds)
{
    // round counters double-buffered by round parity: a round's pair is reset
    // after the round's first barrier, which every thread passes only after
    // reading the pair two rounds back
    __shared__ int s_n2[2], s_nf[2];
    // claims of rounds with at most LCL_MAX_LIST items (fused engines), one
    // table per round parity: the other one is cleared while a round applies
    __shared__ int s_tag[2][LCL_CAP];
    __shared__ unsigned long long s_key[2][LCL_CAP];
    int n = min(cnt[C_NACT], cap);  // A, B and lf hold cap entries each
    unsigned ep = epoch0;
    int r = 0, err = 0, flips = 0;
    int nf = 0;
    if (threadIdx.x == 0) {
        s_n2[0] = 0;
        s_nf[0] = 0;
    }
    // LDS starts with whatever the CU held before: clear tags and keys
    for (int i = threadIdx.x; i < 2 * LCL_CAP; i += TAIL_BS) {
        (&s_tag[0][0])[i] = -1;
        (&s_key[0][0])[i] = 0ull;
    }
    __syncthreads();
    while (n > 0 && r < max_rounds) {
        const int q = r & 1;
        const bool lds = E::kFused && n <= LCL_MAX_LIST;
        const LdsClaims cl{s_tag[q], s_key[q]};
        // fused engines settle the previous round's flips while testing
        for (int i = threadIdx.x; i < n; i += TAIL_BS) {
            if (lds) e.fixtest_l(A[i], ep, cl);
            else if (E::kFused) e.fixtest(A[i], ep, false, false);
            else e.test(A[i], ep, false, false);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s_n2[q ^ 1] = 0;
            s_nf[q ^ 1] = 0;
        }
        for (int i = threadIdx.x; i < LCL_CAP; i += TAIL_BS) {
            s_tag[q ^ 1][i] = -1;
            s_key[q ^ 1][i] = 0ull;
        }
        // block-uniform trip count, so each wave reserves its appends with
        // one LDS atomic per counter
        for (int i0 = 0; i0 < n; i0 += TAIL_BS) {
            const int i = i0 + (int)threadIdx.x;
            int t = -1, nn = -1, res = 0;
            if (i < n) {
                t = A[i];
                res = lds ? e.apply_l(t, ep, &nn, cl) : e.template apply<true>(t, ep, &nn);
            }
            const int b = wave_alloc(res == 1 ? 1 : res == 2 ? 2 : 0, &s_n2[q]);
            const int w = wave_alloc(res == 2 ? 2 : 0, &s_nf[q]);
            if (res == 1) {
                if (b < cap) B[b] = t;
                e.set_mark(t, (int)ep + 1);
            } else if (res == 2) {
                if (b + 1 < cap) {
                    B[b] = t;
                    B[b + 1] = nn;
                }
                if (w + 1 < cap) {
                    lf[w] = t;
                    lf[w + 1] = nn;
                }
            }
        }
        __syncthreads();
        int n2 = s_n2[q];
        nf = s_nf[q];
        if (n2 > cap || nf > cap) {  // cannot happen (one entry per triangle): flagged, not written
            if (threadIdx.x == 0) atomicOr(&cnt[C_ERR], MERR_CAPACITY);
            n = 0;
            break;
        }
        if (!E::kFused) {
            for (int i = threadIdx.x; i < nf; i += TAIL_BS) e.fix(lf[i], ep);
            __syncthreads();
        }
        if (n2 > 0 && nf == 0) {
            err = 1;
            break;
        }
        flips += nf / 2;
        n = n2;
        int* tmp = A;
        A = B;
        B = tmp;
        ep++;
        r++;
    }
    if (E::kFused && n > 0 && nf > 0) {
        // stopped at the round limit with the last round's flips unsettled
        for (int i = threadIdx.x; i < nf; i += TAIL_BS) e.fix(lf[i], ep - 1);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cnt[C_TAIL_N] = n;
        cnt[C_TAIL_R] = r;
        cnt[C_TAIL_ERR] = err;
        cnt[C_TAIL_F] = flips;
    }
}

template <class E>
__global__ __launch_bounds__(kBlock) void k_mark_list(E e, const int* list, int n, int v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        e.set_mark(list ? list[i] : (int)i, v);
}

__global__ __launch_bounds__(kBlock) void k_iota(int* p, int n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = (int)i;
}

__global__ __launch_bounds__(kBlock) void k_pack_faces(const int3* __restrict__ tri, const int3* __restrict__ adj,
                                                       const int* __restrict__ cflag, const int2* __restrict__ pts,
                                                       int T, FRec* __restrict__ r)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (long long)gridDim.x * blockDim.x) {
        FRec x;
        x.v = tri[t];
        x.a = adj[t];
        x.p[0] = pts[x.v.x];
        x.p[1] = pts[x.v.y];
        x.p[2] = pts[x.v.z];
        rec_store(&r[t], x.v, x.a, x.p[0], x.p[1], x.p[2], (cflag ? cflag[t] & 7 : 0) | FREC_NOCAND | FREC_NOPARTNER,
                  0, 0ull);
    }
}

__global__ __launch_bounds__(kBlock) void k_unpack_faces(const FRec* __restrict__ r, int T, int3* __restrict__ tri,
                                                         int3* __restrict__ adj, int* __restrict__ cflag)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (long long)gridDim.x * blockDim.x) {
        const FRec x = rec_load(&r[t]);
        tri[t] = x.v;
        adj[t] = x.a;
        if (cflag) cflag[t] = x.meta & 7;
    }
}

namespace {

// Rounds of test / apply / fix over a triangle list until no edge needs a
// flip.  Round 1 reads `init` (or the implicit list of all T triangles,
// counted in C_CONST); later rounds alternate between two device lists whose
// lengths stay on the device (C_NACT); the host checks round r-1 while round
// r is queued, and lists at or below the tail threshold finish in one
// persistent workgroup.
template <class E>
long long flip_driver(Ctx& c, const E& e, int T, const int* init, int ninit, int* rounds)
{
    DBuf<int> la(T), lb(T), lf(T), lp(T);
    const int* act = init;
    int* nxt = la.p;
    int* other = lb.p;
    if (init) {
        c.set(C_NACT, ninit);
        k_mark_list<E><<<grid_for(ninit), kBlock, 0, c.st>>>(e, init, ninit, (int)c.epoch + 1);
        GCDT_LAUNCH_CHECK();
    } else {
        c.set(C_CONST, T);
    }
    long long flips = 0;
    long long settled = 0;  // fused engines: flipped triangles settled inside test launches
    int r = 0;
    int est = init ? ninit : T;
    bool done = est == 0;
    int checked = 0;  // rounds whose snapshot has been read
    auto check_round = [&](const int* h) {
        if (h[C_ERR]) c.check_err();
        // the highest-keyed proposal always wins both of its triangles
        if (h[C_NACT2] > 0 && h[C_NFLIP] == 0) throw Error(GCDT_ERR_INTERNAL, "flip round without progress");
        flips += h[C_NFLIP] / 2;
        settled += h[C_NFLIP];
        ++checked;
        prof_add_items("flip_test", (checked == 1 && !init) ? h[C_CONST] : h[C_NACT]);
        prof_add_items("flip_apply", h[C_NPROP]);
        if (!E::kFused) prof_add_items("flip_fix", h[C_NFLIP]);  // else settled by the next test launch
    };
    while (!done && est > c.tail_threshold) {
        if (++r > 1000000) throw Error(GCDT_ERR_NOT_CONVERGED, "flip rounds did not converge");
        unsigned ep = c.next_epoch();
        int cnt_idx = (r == 1 && !init) ? C_CONST : C_NACT;
        // begins round r and snapshots round r-1 for the host
        unsigned long long* snap = r > 1 ? c.snap_dst((r - 1) & 1) : nullptr;
        unsigned seq = r > 1 ? c.snap_begin((r - 1) & 1) : 0u;
        k_flip_begin<<<1, C_NUM, 0, c.st>>>(c.dcnt.p, r > 1, snap, seq);
        GCDT_LAUNCH_CHECK();
        with_items(c.flip_items, [&](auto it) {
            constexpr int IT = decltype(it)::value;
            {
                ProfScope ps("flip_test", c.st, 0);
                k_flip_test<E, IT><<<grid_for((est + IT - 1) / IT), kBlock, 0, c.st>>>(
                    e, act, c.dcnt.p + cnt_idx, (init && act == init) ? ninit : T, ep, c.single_tester, lp.p,
                    c.dcnt.p + C_NPROP, T, c.dcnt.p + C_ERR);
                GCDT_LAUNCH_CHECK();
            }
            {
                ProfScope ps("flip_apply", c.st, 0);
                k_flip_apply<E, IT><<<grid_for((est + IT - 1) / IT), kBlock, 0, c.st>>>(
                    e, lp.p, c.dcnt.p + C_NPROP, ep, nxt, c.dcnt.p + C_NACT2, lf.p, c.dcnt.p + C_NFLIP, T,
                    c.dcnt.p + C_ERR);
                GCDT_LAUNCH_CHECK();
            }
        });
        if (!E::kFused) {
            ProfScope ps("flip_fix", c.st, 0);
            k_flip_fix<E><<<grid_for(est), kBlock, 0, c.st>>>(e, lf.p, c.dcnt.p + C_NFLIP, T, ep);
            GCDT_LAUNCH_CHECK();
        }
        act = nxt;
        std::swap(nxt, other);
        if (r >= 2) {
            const int* h = c.wait_snapshot((r - 1) & 1);  // round r-1 (round r is queued)
            check_round(h);
            if (h[C_NACT2] == 0) done = true;
            est = h[C_NACT2];
        }
    }
    if (r >= 1) {
        c.snapshot(r & 1);
        const int* h = c.wait_snapshot(r & 1);
        check_round(h);
        if (h[C_NACT2] == 0) done = true;
        est = h[C_NACT2];
        // the last round's flips are settled by the tail (or the fix below)
        if (E::kFused) prof_add_items("flip_test_fix", settled - h[C_NFLIP]);
        if (E::kFused && done && h[C_NFLIP] > 0) {
            // settle the last round's flips (no test follows; the tail's first
            // test settles them when the tail runs)
            ProfScope ps("flip_fix", c.st, h[C_NFLIP]);
            k_flip_fix<E><<<grid_for(h[C_NFLIP]), kBlock, 0, c.st>>>(e, lf.p, c.dcnt.p + C_NFLIP, T, c.epoch);
            GCDT_LAUNCH_CHECK();
        }
    }
    if (!done) {
        // the rest in one persistent workgroup
        unsigned ep0 = c.epoch + 1;
        if (!act) {  // small mesh: materialise the implicit list
            k_iota<<<grid_for(est), kBlock, 0, c.st>>>(lb.p, est);
            GCDT_LAUNCH_CHECK();
            k_mark_list<E><<<grid_for(est), kBlock, 0, c.st>>>(e, nullptr, est, (int)ep0);
            GCDT_LAUNCH_CHECK();
            act = lb.p;
            nxt = la.p;
            c.set(C_NACT, est);
        } else if (init && act == init) {
            // the tail alternates its two lists: work on a copy of the
            // caller's list (C_NACT holds ninit already)
            GCDT_HIP(hipMemcpyAsync(lb.p, init, sizeof(int) * (size_t)ninit, hipMemcpyDeviceToDevice, c.st));
            act = lb.p;
            nxt = la.p;
        } else if (r >= 1) {
            k_flip_begin<<<1, C_NUM, 0, c.st>>>(c.dcnt.p, 1, nullptr, 0u);
            GCDT_LAUNCH_CHECK();
        }
        {
            ProfScope ps("flip_tail", c.st, est);
            k_flip_tail<E><<<1, TAIL_BS, 0, c.st>>>(e, (int*)act, nxt, lf.p, c.dcnt.p, T, ep0, 1 << 20);
            GCDT_LAUNCH_CHECK();
        }
        c.read(0, C_NUM);
        if (c.hcnt[C_ERR]) c.check_err();
        if (c.hcnt[C_TAIL_ERR]) throw Error(GCDT_ERR_INTERNAL, "flip round without progress (tail)");
        r += c.hcnt[C_TAIL_R];
        c.epoch += (unsigned)c.hcnt[C_TAIL_R];
        flips += c.hcnt[C_TAIL_F];
        if (c.hcnt[C_TAIL_N] != 0) throw Error(GCDT_ERR_NOT_CONVERGED, "flip tail did not converge");
    }
    if (rounds) *rounds = r;
    return flips;
}

}  // namespace

long long flip_rounds_records(Ctx& c, FRec* rec, const Mesh& m, int T, int* rounds)
{
    if (T <= 0) {
        if (rounds) *rounds = 0;
        return 0;
    }
    EngPacked e;
    e.f = FlipP{rec, c.dcnt.p + C_ERR};
    long long flips = flip_driver(c, e, T, nullptr, 0, rounds);
    {
        ProfScope ps("flip_unpack", c.st, T);
        k_unpack_faces<<<grid_for(T), kBlock, 0, c.st>>>(rec, T, m.tri, m.adj, m.cflag);
        GCDT_LAUNCH_CHECK();
    }
    return flips;
}

long long flip_rounds(Ctx& c, const Mesh& m, int T, const int* init, int ninit, int* rounds)
{
    if (T <= 0) {
        if (rounds) *rounds = 0;
        return 0;
    }
    if (!init) {
        // whole mesh: packed records for the rounds, written back at the end
        DBuf<FRec> rec(T);
        {
            ProfScope ps("flip_pack", c.st, T);
            k_pack_faces<<<grid_for(T), kBlock, 0, c.st>>>(m.tri, m.adj, m.cflag, m.pts, T, rec.p);
            GCDT_LAUNCH_CHECK();
        }
        return flip_rounds_records(c, rec.p, m, T, rounds);
    }
    DBuf<unsigned long long> claim(T);
    fill_u64(c, claim.p, 0, T);
    DBuf<int> stamp(T), partner(T), cand(T), mark(T);
    fill_i32(c, stamp.p, 0, T);
    fill_i32(c, mark.p, 0, T);
    EngSoA e;
    e.f.m = m;
    e.f.claim = claim.p;
    e.f.stamp = stamp.p;
    e.f.partner = partner.p;
    e.f.cand = cand.p;
    e.f.mark = mark.p;
    e.f.err = c.dcnt.p + C_ERR;
    return flip_driver(c, e, T, init, ninit, rounds);
}

// ================================================================== batches of instances

// instance of each point (binary search over the point offsets)
__global__ __launch_bounds__(kBlock) void k_batch_vinst(const long long* off, int K, long long NP, int* vinst)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < NP; i += (long long)gridDim.x * blockDim.x) {
        int lo = 0, hi = K;  // last instance with off[inst] <= i
        while (lo + 1 < hi) {
            int mid = (lo + hi) >> 1;
            if (off[mid] <= i) lo = mid;
            else hi = mid;
        }
        vinst[i] = lo;
    }
}

// corners NP+4i..NP+4i+3 and the two triangles 2i, 2i+1 of every instance
__global__ __launch_bounds__(kBlock) void k_batch_init(int2* pts, int3* tri, int3* adj, int K, long long NP)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (long long)gridDim.x * blockDim.x) {
        int c = (int)(NP + 4 * i);
        pts[c] = make_int2(0, 0);
        pts[c + 1] = make_int2(GRID_R, 0);
        pts[c + 2] = make_int2(GRID_R, GRID_R);
        pts[c + 3] = make_int2(0, GRID_R);
        int t = (int)(2 * i);
        tri[t] = make_int3(c, c + 1, c + 3);
        tri[t + 1] = make_int3(c + 1, c + 2, c + 3);
        adj[t] = make_int3(t + 1, -1, -1);
        adj[t + 1] = make_int3(-1, t, -1);
    }
}

// constraint vertex ids: instance-local -> global; ids outside the instance
// become -1 (rejected by the constraint checks)
__global__ __launch_bounds__(kBlock) void k_batch_cons(const int2* in, const long long* coff, const long long* poff, int K,
                                                       long long NC, int2* out)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < NC; i += (long long)gridDim.x * blockDim.x) {
        int lo = 0, hi = K;
        while (lo + 1 < hi) {
            int mid = (lo + hi) >> 1;
            if (coff[mid] <= i) lo = mid;
            else hi = mid;
        }
        long long base = poff[lo], n = poff[lo + 1] - poff[lo];
        int2 e = in[i];
        e.x = (e.x >= 0 && e.x < n) ? (int)(base + e.x) : -1;
        e.y = (e.y >= 0 && e.y < n) ? (int)(base + e.y) : -1;
        out[i] = e;
    }
}

// scatter the global mesh back into per-instance lists with local vertex ids.
// Leaves are in vertex order and point ids instance-major, so an instance's
// triangles fill one range of slots [first, last] (flips and cavities reuse
// slots in place): a triangle's place in its list is its slot minus first,
// the same on every run.  A range that does not hold exactly the instance's
// triangle count flags MERR_INTERNAL_COUNT.
__device__ inline int batch_inst_of(int3 t, const int* vinst, long long NP)
{
    return t.x < NP ? vinst[t.x] : (int)((t.x - NP) >> 2);
}

// the ends of each instance's run of slots, where the instance changes (one
// store per end; a split run stores an end twice and fails the count check)
__global__ __launch_bounds__(kBlock) void k_batch_span(const int3* tri, int T, const int* vinst, long long NP,
                                                       int* first, int* last)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        const int inst = batch_inst_of(tri[i], vinst, NP);
        if (i == 0 || batch_inst_of(tri[i - 1], vinst, NP) != inst) first[inst] = (int)i;
        if (i == T - 1 || batch_inst_of(tri[i + 1], vinst, NP) != inst) last[inst] = (int)i;
    }
}

__global__ __launch_bounds__(kBlock) void k_batch_out(const int3* tri, int T, const int* vinst, const int* perm,
                                                      const long long* poff, const long long* toff, long long NP,
                                                      const int* first, const int* last, int3* out, int* err)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        const int3 t = tri[i];
        const int inst = batch_inst_of(t, vinst, NP);
        const long long base = poff[inst], n = poff[inst + 1] - poff[inst];
        if ((long long)last[inst] - first[inst] + 1 != toff[inst + 1] - toff[inst]) {
            atomicOr(err, MERR_INTERNAL_COUNT);
            continue;
        }
        auto local = [&](int v) -> int {
            return v < NP ? (int)((perm ? perm[v] : v) - base) : (int)(n + ((v - NP) & 3));
        };
        out[toff[inst] + (i - first[inst])] = make_int3(local(t.x), local(t.y), local(t.z));
    }
}

void batch_vinst(Ctx& c, const long long* off, int K, long long NP, int* vinst)
{
    if (NP <= 0) return;
    k_batch_vinst<<<grid_for(NP), kBlock, 0, c.st>>>(off, K, NP, vinst);
    GCDT_LAUNCH_CHECK();
}

// Batch points counting-sorted by instance, then by the Morton order of a
// per-instance grid of 4^L cells (L in [4, 10], about one cell per point, the
// instance's cells numbered from cbase[i]): instance ranges are kept,
// neighbours in the plane become neighbours in memory down to single points,
// and each 16x16 cell of an instance is a contiguous range whose first point
// is a sample.  The same one-atomic counting sort as morton_order: ranks in
// inv[], then an inclusive scan, then a gather of each cell's start.
struct BatchCells {
    const int* vinst;
    const int* cbase;
    const int* lvl;
    __device__ int cell(int i, int2 p) const
    {
        int inst = vinst[i];
        return cbase[inst] + (int)(morton_cell(p) >> (2 * (10 - lvl[inst])));
    }
};

// points by (instance, fine cell), each cell in input order (the count,
// scatter and per-cell reorder of the single solve's Morton order)
__global__ __launch_bounds__(kBlock) void k_batch_cell_count(BatchCells b, const int2* pts, long long NP, int* cnt,
                                                             int* rank, int* huge)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < NP; i += (long long)gridDim.x * blockDim.x) {
        const int r = atomicAdd(&cnt[b.cell((int)i, pts[i])], 1);
        rank[i] = r;
        if (r == CELL_BLOCK_MAX) atomicAdd(huge, 1);
    }
}

__global__ __launch_bounds__(kBlock) void k_batch_cell_scatter(BatchCells b, const int2* pts, long long NP,
                                                               const int* cell_end, const int* rank, int4* eb)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < NP; i += (long long)gridDim.x * blockDim.x) {
        const int2 p = pts[i];
        const int cell = b.cell((int)i, p);
        const int pos = (cell ? cell_end[cell - 1] : 0) + rank[i];
        eb[pos] = make_int4((int)i, cell, p.x, p.y);
    }
}

void batch_order(Ctx& c, const int2* pts, const int* vinst, const int* cbase, const int* lvl, long long NP,
                 long long ncells, int2* out, int* perm, int* inv, int* cell_end)
{
    fill_i32(c, cell_end, 0, ncells + 1);
    if (NP <= 0) return;
    BatchCells b{vinst, cbase, lvl};
    c.set(C_HUGE, 0);
    k_batch_cell_count<<<grid_for(NP), kBlock, 0, c.st>>>(b, pts, NP, cell_end, inv, c.dcnt.p + C_HUGE);
    GCDT_LAUNCH_CHECK();
    scan_incl(c, cell_end, cell_end, ncells);
    DBuf<int4> eb((size_t)NP);
    k_batch_cell_scatter<<<grid_for(NP), kBlock, 0, c.st>>>(b, pts, NP, cell_end, inv, eb.p);
    GCDT_LAUNCH_CHECK();
    cell_reorder(c, CellKeyPoint{pts, perm, out, inv}, cell_end, eb.p, NP, c.get(C_HUGE) > 0);
}

// Sorted range of 16x16 cell c (Morton index) of instance i: the fine cells
// [cbase + (c << 2(L-4)), cbase + ((c+1) << 2(L-4))).
__device__ inline int2 coarse_range(const int* cbase, const int* lvl, const int* cell_end, long long idx)
{
    int inst = (int)(idx >> 8), c = (int)(idx & 255);
    int sh = 2 * (lvl[inst] - 4);
    int f0 = cbase[inst] + (c << sh), f1 = cbase[inst] + ((c + 1) << sh);
    return make_int2(f0 ? cell_end[f0 - 1] : 0, cell_end[f1 - 1]);
}

__global__ __launch_bounds__(kBlock) void k_range_flags(const int* cbase, const int* lvl, const int* cell_end,
                                                        long long ncoarse, int* flags)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < ncoarse; i += (long long)gridDim.x * blockDim.x) {
        int2 r = coarse_range(cbase, lvl, cell_end, i);
        flags[i] = r.y > r.x;
    }
}

__global__ __launch_bounds__(kBlock) void k_range_emit(const int* cbase, const int* lvl, const int* cell_end,
                                                       const int* offs, long long ncoarse, int* ids)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < ncoarse; i += (long long)gridDim.x * blockDim.x) {
        int2 r = coarse_range(cbase, lvl, cell_end, i);
        if (r.y > r.x) ids[offs[i]] = r.x;
    }
}

int batch_sample_sorted(Ctx& c, const int* cbase, const int* lvl, const int* cell_end, int K, int* ids)
{
    long long ncoarse = (long long)K << 8;
    DBuf<int> offs((size_t)ncoarse + 1);
    k_range_flags<<<grid_for(ncoarse), kBlock, 0, c.st>>>(cbase, lvl, cell_end, ncoarse, offs.p);
    GCDT_LAUNCH_CHECK();
    scan_excl(c, offs.p, offs.p, ncoarse);
    k_range_emit<<<grid_for(ncoarse), kBlock, 0, c.st>>>(cbase, lvl, cell_end, offs.p, ncoarse, ids);
    GCDT_LAUNCH_CHECK();
    return c.get(C_TOTAL);
}

// Listed points located in their instance's two seed triangles (the side of
// the anti-diagonal).
__global__ __launch_bounds__(kBlock) void k_batch_locate_ids(const int2* pts, const int* vinst, const int* ids, int n,
                                                             int* loc, int* pend)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        int p = ids[i];
        int t0 = 2 * vinst[p];
        int o = orient2d(make_int2(GRID_R, 0), make_int2(0, GRID_R), pts[p]);
        loc[p] = o > 0 ? t0 * 4 + 3 : (o < 0 ? (t0 + 1) * 4 + 3 : t0 * 4 + 0);
        pend[i] = p;
    }
}

void batch_locate_ids(Ctx& c, const int2* pts, const int* vinst, const int* ids, int n, int* loc, int* pend)
{
    if (n <= 0) return;
    k_batch_locate_ids<<<grid_for(n), kBlock, 0, c.st>>>(pts, vinst, ids, n, loc, pend);
    GCDT_LAUNCH_CHECK();
}

void batch_init(Ctx& c, int2* pts, int3* tri, int3* adj, int K, long long NP)
{
    k_batch_init<<<grid_for(K), kBlock, 0, c.st>>>(pts, tri, adj, K, NP);
    GCDT_LAUNCH_CHECK();
}

void batch_cons(Ctx& c, const int2* in, const long long* coff, const long long* poff, int K, long long NC, int2* out)
{
    if (NC <= 0) return;
    k_batch_cons<<<grid_for(NC), kBlock, 0, c.st>>>(in, coff, poff, K, NC, out);
    GCDT_LAUNCH_CHECK();
}

void batch_out(Ctx& c, const int3* tri, int T, const int* vinst, const int* perm, const long long* poff,
               const long long* toff, long long NP, int K, int3* out)
{
    DBuf<int> first(K), last(K);
    fill_i32(c, first.p, INT_MAX, K);
    fill_i32(c, last.p, -1, K);
    k_batch_span<<<grid_for(T), kBlock, 0, c.st>>>(tri, T, vinst, NP, first.p, last.p);
    GCDT_LAUNCH_CHECK();
    k_batch_out<<<grid_for(T), kBlock, 0, c.st>>>(tri, T, vinst, perm, poff, toff, NP, first.p, last.p, out,
                                                  c.dcnt.p + C_ERR);
    GCDT_LAUNCH_CHECK();
    c.check_err();
}

// ================================================================== flags -> list

__global__ __launch_bounds__(kBlock) void k_flag_scatter(const int* flags, const int* off, int n, int* list)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        if (flags[i]) list[off[i]] = (int)i;
}

int flags_to_list(Ctx& c, const int* flags, int n, int* list)
{
    DBuf<int> off((size_t)n + 1);
    scan_excl(c, flags, off.p, n);
    k_flag_scatter<<<grid_for(n), kBlock, 0, c.st>>>(flags, off.p, n, list);
    GCDT_LAUNCH_CHECK();
    return c.get(C_TOTAL);
}

// ================================================================== phase check

// Per-triangle checks of a mesh between phases (the "validate" tunable):
// bits of C_VAL name what failed, C_VAL_FIRST the lowest failing triangle.
enum : int { VAL_IDS = 1, VAL_CW = 2, VAL_ADJ = 4, VAL_CFLAG = 8, VAL_DELAUNAY = 16 };

__global__ __launch_bounds__(kBlock) void k_validate_mesh(const int2* pts, long long nv, const int3* tri,
                                                          const int3* adj, const int* cflag, int T, int delaunay,
                                                          int* bad, int* first)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        const int3 v = tri[i];
        int f = 0;
        if (v.x < 0 || v.y < 0 || v.z < 0 || v.x >= nv || v.y >= nv || v.z >= nv) {
            f = VAL_IDS;
        } else {
            const int2 A = pts[v.x], B = pts[v.y], C = pts[v.z];
            if (orient2d(A, B, C) <= 0) f |= VAL_CW;
            const int3 a = adj[i];
            const int cf = cflag ? cflag[i] : 0;
            for (int k = 0; k < 3; k++) {
                const int n = get3(a, k);
                if (n < 0) continue;  // the bounding square's sides
                if (n >= T) {
                    f |= VAL_ADJ;
                    continue;
                }
                const int u = get3(v, nx3(k)), w = get3(v, pv3(k));
                const int3 nt = tri[n];
                if (!has2(nt, u, w)) {
                    f |= VAL_ADJ;
                    continue;
                }
                const int j = slot_other(nt, u, w);
                if (get3(adj[n], j) != (int)i) {
                    f |= VAL_ADJ;
                    continue;
                }
                const int ck = (cf >> k) & 1;
                if (cflag && ck != ((cflag[n] >> j) & 1)) f |= VAL_CFLAG;
                const int o = get3(nt, j);
                if (delaunay && !ck && n > i && (o < 0 || o >= nv || incircle(A, B, C, pts[o]) > 0)) f |= VAL_DELAUNAY;
            }
        }
        if (f) {
            atomicOr(bad, f);
            atomicMin(first, (int)i);
        }
    }
}

void check_mesh(Ctx& c, const int2* pts, long long nv, const int3* tri, const int3* adj, const int* cflag, int T,
                bool delaunay, int* failed, int* first)
{
    *failed = 0;
    *first = -1;
    if (T <= 0) return;
    c.set(C_VAL, 0);
    c.set(C_VAL_FIRST, INT_MAX);
    k_validate_mesh<<<grid_for(T), kBlock, 0, c.st>>>(pts, nv, tri, adj, cflag, T, delaunay ? 1 : 0, c.dcnt.p + C_VAL,
                                                      c.dcnt.p + C_VAL_FIRST);
    GCDT_LAUNCH_CHECK();
    c.read(C_VAL, 2);
    *failed = c.hcnt[C_VAL];
    if (*failed) *first = c.hcnt[C_VAL_FIRST];
}

void validate_mesh(Ctx& c, const char* phase, const Mesh& m, long long nv, int T, bool delaunay)
{
    if (!Tunables::get().validate || T <= 0) return;
    int f = 0, first = -1;
    check_mesh(c, m.pts, nv, m.tri, m.adj, m.cflag, T, delaunay, &f, &first);
    if (!f) return;
    std::string what;
    const char* names[] = {"vertex ids", "orientation", "adjacency", "constraint bits", "local Delaunay"};
    for (int b = 0; b < 5; b++)
        if (f & (1 << b)) what += std::string(what.empty() ? "" : ", ") + names[b];
    throw Error(GCDT_ERR_INTERNAL, std::string("mesh check after ") + phase + " failed (" + what + "), first at triangle " +
                                       std::to_string(first));
}

// ================================================================== leaves

__global__ __launch_bounds__(kBlock) void k_leaf_flags(const int* sons, int T, int* flags)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x)
        flags[i] = sons[i] == -1;
}

__global__ __launch_bounds__(kBlock) void k_leaf_scatter(const int3* tri, const int3* adj, const int* sons, const int* off,
                                                         int T, int3* tri2, int3* adj2, long long cap, int* err)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        if (sons[i] != -1) continue;
        int o = off[i];
        if (o >= cap) {
            atomicOr(err, MERR_CAPACITY);
            continue;
        }
        tri2[o] = tri[i];
        if (adj && adj2) {
            int3 a = adj[i];
            a.x = a.x >= 0 ? off[a.x] : -1;
            a.y = a.y >= 0 ? off[a.y] : -1;
            a.z = a.z >= 0 ? off[a.z] : -1;
            adj2[o] = a;
        }
    }
}

// Leaves in Morton order of their centroid's 1024^2 cell (the reference's
// reorderFaces, kernel.cu:1886-1964, done as a counting sort fused with the
// compaction): neighbouring triangles end up near each other in memory.
// Sort key of a leaf for the compacted mesh order: the Morton cell of its
// centroid on the 2048x2048 grid over the square (about five leaves per cell
// at 10M points, so nearly every cell is ranked lane by lane).
#ifndef GCDT_LEAF_CELL_BITS
#define GCDT_LEAF_CELL_BITS 11
#endif
constexpr int LEAF_CELL_BITS = GCDT_LEAF_CELL_BITS;
constexpr long long LEAF_CELLS = 1LL << (2 * LEAF_CELL_BITS);
__device__ inline unsigned morton_cell_leaf(int2 p)
{
    unsigned x = (unsigned)p.x >> (30 - LEAF_CELL_BITS), y = (unsigned)p.y >> (30 - LEAF_CELL_BITS);
    unsigned r = 0;
    for (int b = 0; b < LEAF_CELL_BITS; b++) r |= (((x >> b) & 1u) << (2 * b)) | (((y >> b) & 1u) << (2 * b + 1));
    return r;
}
struct LeafKeyOne {
    __device__ int key(const int3&, int2 g) const { return (int)morton_cell_leaf(g); }
};

// The count pass keeps each leaf's rank within its cell in rank[]; the dest
// pass adds the cell's start to it without a second atomic.
template <class KEY>
__global__ __launch_bounds__(kBlock) void k_leaf_cells(KEY kf, const int2* pts, const int3* tri, const int* sons, int T,
                                                       int* cnt, int* cellof, int* rank, int* huge)
{
    WAVE_LOOP(i, T)
    {
        int cell = -1;
        if (i < T && sons[i] == -1) {
            int3 t = tri[i];
            int2 a = pts[t.x], b = pts[t.y], c = pts[t.z];
            int2 g = make_int2((int)(((long long)a.x + b.x + c.x) / 3), (int)(((long long)a.y + b.y + c.y) / 3));
            cell = kf.key(t, g);
        }
        // consecutive leaves were allocated together and mostly share a cell
        int r = wave_run_add(cnt, cell);
        if (cell >= 0) {
            cellof[i] = cell;
            rank[i] = r;
            if (r == CELL_BLOCK_MAX) atomicAdd(huge, 1);
        }
    }
}

// dest[] holds the ranks on entry and the arrival-order destinations on exit
// (-1 for dead triangles); eb[] gets (leaf, cell) at each destination.  The
// per-cell reorder then writes the final destinations and their inverse.
__global__ __launch_bounds__(kBlock) void k_leaf_dest(const int* sons, const int* cellof, int T, const int* start,
                                                      int* dest, int2* eb, long long cap)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        int d = sons[i] == -1 ? start[cellof[i]] + dest[i] : -1;
        dest[i] = d;
        if (d >= 0 && d < cap) eb[d] = make_int2((int)i, cellof[i]);
    }
}

// Where the moved leaves go: plain tri/adj arrays, or straight into the
// packed flip records (coordinates gathered here, so the flips need no
// separate pack pass over a freshly written mesh).
struct LeafToArrays {
    int3* tri2;
    int3* adj2;
    __device__ void put(long long o, int3 v, int3 a) const
    {
        tri2[o] = v;
        adj2[o] = a;
    }
};

struct LeafToRecords {
    FRec* rec;
    const int2* pts;
    __device__ void put(long long o, int3 v, int3 a) const
    {
        rec_store(&rec[o], v, a, pts[v.x], pts[v.y], pts[v.z], FREC_NOCAND | FREC_NOPARTNER, 0, 0ull);
    }
};

// Walks the destinations in order and gathers each one's leaf, so the
// (64-byte record) writes are contiguous and the reads do the scattering.
template <class OUT>
__global__ __launch_bounds__(kBlock) void k_leaf_move(const int3* tri, const int3* adj, const int* dest,
                                                      const int* src, int n, OUT out)
{
    for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x) {
        int i = src[o];
        int3 a = adj[i];
        a.x = a.x >= 0 ? dest[a.x] : -1;
        a.y = a.y >= 0 ? dest[a.y] : -1;
        a.z = a.z >= 0 ? dest[a.z] : -1;
        out.put(o, tri[i], a);
    }
}

template <class KEY, class OUT>
static int filter_leaves_sorted(Ctx& c, const KEY& kf, long long ncells, const int2* pts, const int3* tri,
                                const int3* adj, const int* sons, int T, const OUT& out, long long cap)
{
    DBuf<int> cnt((size_t)ncells + 1), cellof(T > 0 ? T : 1), dest(T > 0 ? T : 1);
    fill_i32(c, cnt.p, 0, ncells + 1);
    c.set(C_HUGE, 0);
    k_leaf_cells<KEY><<<grid_for(T), kBlock, 0, c.st>>>(kf, pts, tri, sons, T, cnt.p, cellof.p, dest.p,
                                                        c.dcnt.p + C_HUGE);
    GCDT_LAUNCH_CHECK();
    scan_excl_total(c, cnt.p, cnt.p, ncells);
    c.read(C_TOTAL, C_HUGE - C_TOTAL + 1);
    const int n = c.hcnt[C_TOTAL];
    const bool huge = c.hcnt[C_HUGE] > 0;
    if (n > cap) throw Error(GCDT_ERR_CAPACITY, "more leaves than the output holds");
    DBuf<int> src(n > 0 ? n : 1);
    {
        DBuf<int2> eb(n > 0 ? n : 1);
        k_leaf_dest<<<grid_for(T), kBlock, 0, c.st>>>(sons, cellof.p, T, cnt.p, dest.p, eb.p, cap);
        GCDT_LAUNCH_CHECK();
        cellof.release();
        // each cell's leaves by their vertex ids (the count pass placed them
        // in arrival order); cnt holds the cells' starts, cnt[c + 1] their ends
        cell_reorder(c, CellKeyTri{tri, src.p, dest.p}, cnt.p + 1, eb.p, n, huge);
    }
    if (n > 0) {
        ProfScope ps("leaf_move", c.st, n);
        k_leaf_move<OUT><<<grid_for(n), kBlock, 0, c.st>>>(tri, adj, dest.p, src.p, n, out);
        GCDT_LAUNCH_CHECK();
    }
    return n;
}

int filter_leaves_morton(Ctx& c, const int2* pts, const int3* tri, const int3* adj, const int* sons, int T, int3* tri2,
                         int3* adj2, long long cap)
{
    return filter_leaves_sorted(c, LeafKeyOne{}, LEAF_CELLS, pts, tri, adj, sons, T, LeafToArrays{tri2, adj2}, cap);
}

int filter_leaves_morton_rec(Ctx& c, const int2* pts, const int3* tri, const int3* adj, const int* sons, int T,
                             FRec* rec, long long cap)
{
    return filter_leaves_sorted(c, LeafKeyOne{}, LEAF_CELLS, pts, tri, adj, sons, T, LeafToRecords{rec, pts}, cap);
}

// Batch meshes: instance-major, then the Morton order of the centroid's cell
// in the instance's own grid of 4^L cells (the point order's cells, about one
// point each), so a cell holds a few leaves and its reorder stays in one wave.
struct LeafKeyInst {
    const int* vinst;
    const int* cbase;
    const int* lvl;
    long long NP;  // corners of instance i are NP + 4i .. NP + 4i + 3
    __device__ int key(const int3& t, int2 g) const
    {
        const int inst = t.x < NP ? vinst[t.x] : (int)((t.x - NP) >> 2);
        return cbase[inst] + (int)(morton_cell(g) >> (2 * (10 - lvl[inst])));
    }
};

int filter_leaves_inst(Ctx& c, const int2* pts, const int* vinst, const int* cbase, const int* lvl, long long ncells,
                       long long NP, const int3* tri, const int3* adj, const int* sons, int T, int3* tri2, int3* adj2,
                       long long cap)
{
    return filter_leaves_sorted(c, LeafKeyInst{vinst, cbase, lvl, NP}, ncells, pts, tri, adj, sons, T,
                                LeafToArrays{tri2, adj2}, cap);
}

int filter_leaves_inst_rec(Ctx& c, const int2* pts, const int* vinst, const int* cbase, const int* lvl,
                           long long ncells, long long NP, const int3* tri, const int3* adj, const int* sons, int T,
                           FRec* rec, long long cap)
{
    return filter_leaves_sorted(c, LeafKeyInst{vinst, cbase, lvl, NP}, ncells, pts, tri, adj, sons, T,
                                LeafToRecords{rec, pts}, cap);
}

int filter_leaves(Ctx& c, const int3* tri, const int3* adj, const int* sons, int T, int3* tri2, int3* adj2,
                  long long cap)
{
    DBuf<int> off((size_t)T + 1);
    k_leaf_flags<<<grid_for(T), kBlock, 0, c.st>>>(sons, T, off.p);
    GCDT_LAUNCH_CHECK();
    scan_excl(c, off.p, off.p, T);
    k_leaf_scatter<<<grid_for(T), kBlock, 0, c.st>>>(tri, adj, sons, off.p, T, tri2, adj2, cap, c.dcnt.p + C_ERR);
    GCDT_LAUNCH_CHECK();
    int n = c.get(C_TOTAL);
    c.check_err();
    return n;
}

// ================================================================== bitmap

template <int LOG>
__device__ inline void tri_cells(int2 A, int2 B, int2 C, int& x0, int& y0, int& x1, int& y1)
{
    constexpr int sh = 30 - LOG;
    x0 = min(A.x, min(B.x, C.x)) >> sh;
    x1 = max(A.x, max(B.x, C.x)) >> sh;
    y0 = min(A.y, min(B.y, C.y)) >> sh;
    y1 = max(A.y, max(B.y, C.y)) >> sh;
}

// Closed triangle vs closed box, exact (separating-axis test: the box axes,
// then each triangle edge with all four corners strictly outside).
// Same predicate as the reference's intersect(), boostmap.cu:30-66.
__device__ inline bool tri_meets_box(int2 A, int2 B, int2 C, int bx0, int by0, int bx1, int by1)
{
    if (max(A.x, max(B.x, C.x)) < bx0 || min(A.x, min(B.x, C.x)) > bx1) return false;
    if (max(A.y, max(B.y, C.y)) < by0 || min(A.y, min(B.y, C.y)) > by1) return false;
    int2 q[4] = {make_int2(bx0, by0), make_int2(bx1, by0), make_int2(bx1, by1), make_int2(bx0, by1)};
    int2 P[3] = {A, B, C};
#pragma unroll
    for (int e = 0; e < 3; e++) {
        int2 u = P[e], v = P[(e + 1) % 3];
        bool out = true;
#pragma unroll
        for (int j = 0; j < 4; j++) out &= orient2d(u, v, q[j]) < 0;
        if (out) return false;
    }
    return true;
}

template <class G>
__global__ __launch_bounds__(kBlock) void k_bm_count(G g, const int2* pts, const int3* tri, int T, int* cnt)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        int3 t = tri[i];
        int x0, y0, x1, y1;
        tri_cells<G::LOG>(pts[t.x], pts[t.y], pts[t.z], x0, y0, x1, y1);
        cnt[i] = (x1 - x0 + 1) * (y1 - y0 + 1);
    }
}

__device__ inline int upper_bound_i(const int* a, int n, long long x)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <class G>
__global__ __launch_bounds__(kBlock) void k_bm_test(G g, const int2* pts, const int3* tri, const int* off, int T,
                                                    long long npairs, int* cellcnt, int* ord, int* cellof)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < npairs; i += (long long)gridDim.x * blockDim.x) {
        int t = upper_bound_i(off, T + 1, i) - 1;
        int local = (int)(i - off[t]);
        int3 tv = tri[t];
        int2 A = pts[tv.x], B = pts[tv.y], C = pts[tv.z];
        int x0, y0, x1, y1;
        tri_cells<G::LOG>(A, B, C, x0, y0, x1, y1);
        int h = y1 - y0 + 1;
        int cx = x0 + local / h, cy = y0 + local % h;
        constexpr int sh = 30 - G::LOG;
        bool hit = tri_meets_box(A, B, C, cx << sh, cy << sh, ((cx + 1) << sh) - 1, ((cy + 1) << sh) - 1);
        int cell = (int)(g.tri_base(tv) + ((long long)cx << G::LOG) + cy);
        cellof[i] = cell;
        ord[i] = hit ? atomicAdd(&cellcnt[cell], 1) : -1;
    }
}

__global__ __launch_bounds__(kBlock) void k_bm_fill(const int* off, int T, long long npairs, const int* ord,
                                                    const int* cellof, const int* cell_off, int* bm)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < npairs; i += (long long)gridDim.x * blockDim.x) {
        int o = ord[i];
        if (o < 0) continue;
        int t = upper_bound_i(off, T + 1, i) - 1;
        bm[cell_off[cellof[i]] + o] = t;
    }
}

template <class G>
static void build_bitmap_g(Ctx& c, const G& g, long long ncells, const int2* pts, const int3* tri, int T,
                           int** offsets, int** tris, int* ntris)
{
    DBuf<int> off((size_t)T + 1);
    k_bm_count<G><<<grid_for(T), kBlock, 0, c.st>>>(g, pts, tri, T, off.p);
    GCDT_LAUNCH_CHECK();
    scan_excl_total(c, off.p, off.p, T);
    long long npairs = c.get(C_TOTAL);
    DBuf<int> ord(npairs), cellof(npairs);
    DBuf<int> cell_off((size_t)ncells + 1);
    fill_i32(c, cell_off.p, 0, ncells + 1);
    if (npairs > 0) {
        k_bm_test<G><<<grid_for(npairs), kBlock, 0, c.st>>>(g, pts, tri, off.p, T, npairs, cell_off.p, ord.p,
                                                            cellof.p);
        GCDT_LAUNCH_CHECK();
    }
    scan_excl_total(c, cell_off.p, cell_off.p, ncells);
    int total = c.get(C_TOTAL);
    DBuf<int> bm(total > 0 ? total : 1);
    if (npairs > 0) {
        k_bm_fill<<<grid_for(npairs), kBlock, 0, c.st>>>(off.p, T, npairs, ord.p, cellof.p, cell_off.p, bm.p);
        GCDT_LAUNCH_CHECK();
    }
    // hand the buffers to the caller; they are released with the sizes
    // (ncells + 1) ints and `total` ints
    *offsets = cell_off.detach();
    *tris = bm.detach();
    *ntris = total;
}

void build_bitmap(Ctx& c, const int2* pts, const int3* tri, int T, int** offsets, int** tris, int* ntris)
{
    build_bitmap_g(c, GridOne{}, BITMAP_CELLS, pts, tri, T, offsets, tris, ntris);
}

void build_bitmap_batch(Ctx& c, const int2* pts, const int* vinst, long long NP, int K, const int3* tri, int T,
                        int** offsets, int** tris, int* ntris)
{
    build_bitmap_g(c, GridInst{vinst, NP}, (long long)K << (2 * GridInst::LOG), pts, tri, T, offsets, tris, ntris);
}

// ================================================================== adjacency from a triangle list

__global__ __launch_bounds__(kBlock) void k_adj_count(const int3* tri, int T, int* cnt)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < 3LL * T; i += (long long)gridDim.x * blockDim.x) {
        int t = (int)(i / 3), k = (int)(i % 3);
        int3 tv = tri[t];
        int u = get3(tv, nx3(k)), v = get3(tv, pv3(k));
        atomicAdd(&cnt[min(u, v)], 1);
    }
}

__global__ __launch_bounds__(kBlock) void k_adj_fill(const int3* tri, int T, const int* off, int* cur, int* bucket)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < 3LL * T; i += (long long)gridDim.x * blockDim.x) {
        int t = (int)(i / 3), k = (int)(i % 3);
        int3 tv = tri[t];
        int u = get3(tv, nx3(k)), v = get3(tv, pv3(k));
        int key = min(u, v);
        bucket[off[key] + atomicAdd(&cur[key], 1)] = (int)i;
    }
}

__global__ __launch_bounds__(kBlock) void k_adj_match(const int3* tri, int T, const int* off, const int* bucket, int* adj)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < 3LL * T; i += (long long)gridDim.x * blockDim.x) {
        int t = (int)(i / 3), k = (int)(i % 3);
        int3 tv = tri[t];
        int u = get3(tv, nx3(k)), v = get3(tv, pv3(k));
        int key = min(u, v);
        int r = -1;
        for (int j = off[key], e = off[key + 1]; j < e; j++) {
            int h = bucket[j];
            if (h == (int)i) continue;
            int3 w = tri[h / 3];
            int kk = h % 3;
            if (get3(w, nx3(kk)) == v && get3(w, pv3(kk)) == u) {
                r = h / 3;
                break;
            }
        }
        adj[i] = r;
    }
}

__global__ __launch_bounds__(kBlock) void k_cons_mark(const int3* tri, const int* off, const int* bucket, const int2* cons,
                                                      int C, int* cflag)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < C; i += (long long)gridDim.x * blockDim.x) {
        int2 e = cons[i];
        int key = min(e.x, e.y), other = max(e.x, e.y);
        for (int j = off[key], f = off[key + 1]; j < f; j++) {
            int h = bucket[j];
            int3 w = tri[h / 3];
            int kk = h % 3;
            int a = get3(w, nx3(kk)), b = get3(w, pv3(kk));
            if (max(a, b) == other) atomicOr(&cflag[h / 3], 1 << kk);
        }
    }
}

void build_adjacency(Ctx& c, const int3* tri, int T, int nv, int3* adj)
{
    DBuf<int> off((size_t)nv + 1), cur(nv), bucket(3LL * T);
    fill_i32(c, off.p, 0, nv + 1);
    fill_i32(c, cur.p, 0, nv);
    k_adj_count<<<grid_for(3LL * T), kBlock, 0, c.st>>>(tri, T, off.p);
    GCDT_LAUNCH_CHECK();
    scan_excl_total(c, off.p, off.p, nv);
    k_adj_fill<<<grid_for(3LL * T), kBlock, 0, c.st>>>(tri, T, off.p, cur.p, bucket.p);
    GCDT_LAUNCH_CHECK();
    k_adj_match<<<grid_for(3LL * T), kBlock, 0, c.st>>>(tri, T, off.p, bucket.p, (int*)adj);
    GCDT_LAUNCH_CHECK();
}

void mark_constraint_edges(Ctx& c, const int3* tri, const int3* adj, int T, int nv, const int2* cons, int C, int* cflag)
{
    (void)adj;
    DBuf<int> off((size_t)nv + 1), cur(nv), bucket(3LL * T);
    fill_i32(c, off.p, 0, nv + 1);
    fill_i32(c, cur.p, 0, nv);
    k_adj_count<<<grid_for(3LL * T), kBlock, 0, c.st>>>(tri, T, off.p);
    GCDT_LAUNCH_CHECK();
    scan_excl_total(c, off.p, off.p, nv);
    k_adj_fill<<<grid_for(3LL * T), kBlock, 0, c.st>>>(tri, T, off.p, cur.p, bucket.p);
    GCDT_LAUNCH_CHECK();
    if (C > 0) {
        k_cons_mark<<<grid_for(C), kBlock, 0, c.st>>>(tri, off.p, bucket.p, cons, C, cflag);
        GCDT_LAUNCH_CHECK();
    }
}

}  // namespace gcdt
