// Host-side interface of the mesh phases (internal to the library).
#pragma once
#include <utility>
#include <vector>

#include "../../include/MemCache.h"
#include "mesh_bodies.hpp"

namespace gcdt {

// Pool blocks released while a solve on a caller's stream may still have
// kernels queued are parked here (per thread, while that solve's Ctx lives)
// and returned to the pool once the stream has drained: a solve on another
// stream could otherwise be handed a block that queued kernels still use.
// Solves on the legacy null stream release at once -- that stream orders
// every later use of the block.
struct PendingFree {
    void* p;
    size_t bytes;
};
extern thread_local std::vector<PendingFree>* g_defer;

// A region of device memory that this thread's device buffers come from
// instead of the pool while it is set (g_arena): the allocation-free solve
// carves every buffer out of a fixed scratch array.  Buffers are taken from
// the top, 256-byte aligned; a released buffer is given back when it is the
// top one (with any released buffers it uncovers), otherwise it is remembered
// and given back once the buffers above it are.  Taking more than the region
// holds throws GCDT_ERR_SCRATCH.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t top = 0;
    size_t high = 0;  // most bytes in use at once
    static constexpr int kFreed = 64;
    size_t freed_off[kFreed];
    size_t freed_end[kFreed];
    int nfreed = 0;
    Arena(void* b, size_t c) : base((char*)b), cap(c) {}
    bool owns(const void* p) const { return (const char*)p >= base && (const char*)p < base + cap; }
    void* take(size_t bytes);
    void give(void* p, size_t bytes);
};
extern thread_local Arena* g_arena;

// device memory for a buffer: the arena when one is set, else the pool
void* dev_take(size_t bytes);
// give back memory from dev_take (or from the pool: the same rule)
void dev_give(void* p, size_t bytes);

// Device buffer from the MemCache pool (or the thread's arena), released on
// destruction.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    void alloc(size_t count)
    {
        release();
        n = count;
        p = (T*)dev_take(sizeof(T) * (count ? count : 1));
    }
    void ensure(size_t count)
    {
        if (count > n || !p) alloc(count);
    }
    void release()
    {
        if (p) dev_give(p, sizeof(T) * (n ? n : 1));
        p = nullptr;
        n = 0;
    }
    T* detach()
    {
        T* q = p;
        p = nullptr;
        n = 0;
        return q;
    }
    ~DBuf() { release(); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
};

// device counters (indices into Ctx::dcnt)
enum : int {
    C_NPEND = 0,
    C_NPEND2 = 1,
    C_NTRI = 2,
    C_TBASE = 3,
    C_ERR = 4,
    C_NACT = 5,
    C_NACT2 = 6,
    C_NFLIP = 7,
    C_TOTAL = 8,   // scan totals
    C_CONST = 9,   // a constant count for implicit lists
    C_NPIECE = 10,
    C_NQ = 11,
    C_NQ2 = 12,
    C_SCR = 13,    // cavity scratch cursor
    C_NREW = 14,   // rewritten triangles this round
    C_NWIN = 15,
    C_TAIL_N = 16,    // persistent tail: remaining list length
    C_TAIL_R = 17,    // rounds done
    C_TAIL_ERR = 18,  // 1 = a round without progress
    C_TAIL_PAR = 19,  // parity: the final list is the second buffer
    C_TAIL_F = 20,    // flips done
    C_NPROP = 21,     // proposals of the current flip round
    C_NW = 22,        // insertion: triangles claimed for the current round
    C_NW2 = 23,       // insertion: triangles claimed for the next round
    C_STN_T = 24,     // crossing pass: triangles in use
    C_STN_S = 25,     // crossing pass: crossing vertices
    C_HUGE = 26,      // count passes: cells of more elements than one workgroup sorts
    C_VAL = 27,       // mesh check: failed checks (VAL_* bits)
    C_VAL_FIRST = 28, // mesh check: first failing triangle
    C_NUM = 32
};

// Process-wide defaults of the round tunables (environment variables
// GCDT_TAIL_THRESHOLD, GCDT_PIVOT, GCDT_SINGLE_TESTER and GCDT_GRID_CAP
// override them at start; a value outside tunable_cleared() makes every
// solve fail with GCDT_ERR_INVALID instead of running with it).
struct Tunables {
    int tail_threshold = 4096;
    int pivot = 1;
    int single_tester = 1;
    int flip_items = 1;  // items per thread of flip test/apply: 1, 2 or 4
    int ins_items = 1;   // items per thread of insertion split/relocate: 1, 2 or 4
    int validate = 0;    // 1: check the mesh after every phase (validate_mesh)
    std::string env_error;
    static Tunables& get();
};
// whether a GPU parity test covers this value of the tunable
bool tunable_cleared(const std::string& name, int value);

// Pinned host pages of a context's counters: the mirror for synchronous
// reads and the mapped ring of pipelined snapshots.
struct HostSlot {
    int* hcnt = nullptr;
    unsigned long long* hring = nullptr;
    unsigned long long* hring_dev = nullptr;
};
// a slot from the process-wide free list, or newly allocated pinned pages
HostSlot take_host_slot();

struct Ctx {
    hipStream_t st = nullptr;
    DBuf<int> dcnt;
    int* hcnt = nullptr;  // pinned mirror
    unsigned epoch = 0;
    DBuf<int> scan_sums;
    // optional caller scratch (the reference's cubTempStorage): scans use it
    // instead of the pool when set, and fail with GCDT_ERR_SCRATCH if small
    int* ext_scratch = nullptr;
    size_t ext_scratch_bytes = 0;
    // lists at or below this length finish in one persistent workgroup
    int tail_threshold = 4096;
    // tunables (environment overrides for experiments): insertion pivot policy
    // and whether both-active edges are tested by the lower triangle only
    int pivot = 1;
    int single_tester = 1;
    int flip_items = 1;
    int ins_items = 1;
    // fixed: use these pinned pages (kept by the caller) instead of a slot
    // from the free list
    explicit Ctx(hipStream_t s, const HostSlot* fixed = nullptr);
    bool fixed_slot = false;
    ~Ctx();
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;
    unsigned next_epoch() { return ++epoch; }
    // copy counters [i, i+k) to hcnt and wait
    void read(int i, int k = 1);
    int get(int i)
    {
        read(i, 1);
        return hcnt[i];
    }
    void set(int i, int v);
    // pipelined snapshots of all counters: a ring of 2 slots in mapped pinned
    // memory, each counter one 64-bit word (sequence number << 32 | value)
    // that a kernel writes with one system-scope store; the host polls a slot
    // until every word carries the slot's current sequence number (no event
    // per round: an event marker between two kernels costs a gap of about
    // 10 us on the device)
    unsigned long long* hring = nullptr;
    unsigned long long* hring_dev = nullptr;  // the same ring, as the device writes it
    unsigned ring_seq[2] = {0, 0};
    int hsnap[2][C_NUM];
    void snapshot(int slot);             // a copy kernel into the slot
    const int* wait_snapshot(int slot);  // polls until the slot is complete
    // a kernel that snapshots the counters itself writes snap_dst(slot) with
    // the sequence number snap_begin(slot) returns
    unsigned long long* snap_dst(int slot) const { return hring_dev + C_NUM * slot; }
    unsigned snap_begin(int slot);
    void zero_counters();
    void check_err();  // throws on a device error flag
    // deferred pool releases of a solve on a non-null stream
    std::vector<PendingFree> deferred;
    std::vector<PendingFree>* prev_defer = nullptr;
};

// Live per-kernel timing (enabled through the C ABI): launches of the
// kernels named in prof_kernels are bracketed by hipEvents on their stream;
// items = list entries the launch processed (for algorithmic bytes).
struct ProfScope {
    int slot = -1;
    hipStream_t st = nullptr;
    ProfScope(const char* name, hipStream_t s, long long items);
    ~ProfScope();
};

void prof_enable(bool on);
void prof_only(const char* name);  // restrict timing to one kernel (NULL: all)
// items processed by a timed kernel, known only after its launch (exact list
// lengths read back from the round counters)
void prof_add_items(const char* name, long long items);
void prof_reset();
bool prof_read(const char* name, double* ms, long long* launches, long long* items);
int prof_names(char* buf, int cap);

// primitives
void fill_i32(Ctx& c, int* p, int v, long long n);
void fill_u64(Ctx& c, unsigned long long* p, unsigned long long v, long long n);
// exclusive scan of n ints (out may alias in); total also to dcnt[C_TOTAL]
void scan_excl(Ctx& c, const int* in, int* out, long long n);
// inclusive scan: out[i] = in[0] + ... + in[i]
void scan_incl(Ctx& c, const int* in, int* out, long long n);
// same, and out[n] = total (out holds n+1 ints)
void scan_excl_total(Ctx& c, const int* in, int* out, long long n);
long long scan_scratch_ints(long long n);

// mesh construction
void init_square(Ctx& c, int2* pts, int3* tri, int3* adj, int n);
// flags GCDT_ERR_INVALID unless every point is strictly inside the square
void check_points(Ctx& c, const int2* pts, int n);
// sample: one point per occupied 256x256 cell (smallest id); returns count
int sample_points(Ctx& c, const int2* pts, int n, int* sample_ids, int2* sample_pts);
// insertion rounds; pend holds npend point ids with loc[] set; tri/adj/sons
// hold ntri slots. Returns the new slot count.
int insert_rounds(Ctx& c, const Ins& s, DBuf<int>& pend, DBuf<int>& pend2, int npend, int ntri,
                  int* rounds);
// flip to (constrained) Delaunay; init list (nullptr = all T triangles)
long long flip_rounds(Ctx& c, const Mesh& m, int T, const int* init, int ninit, int* rounds);
// With the "validate" tunable on: checks m's T triangles over nv vertices
// (counter-clockwise, symmetric adjacency, equal constraint bits on both
// sides, and with delaunay every unconstrained edge locally Delaunay) and
// fails with GCDT_ERR_INTERNAL naming `phase`.  Off: returns at once.
void validate_mesh(Ctx& c, const char* phase, const Mesh& m, long long nv, int T, bool delaunay);
// The checks themselves: *failed gets the failed-check bits (VAL_*), *first the
// lowest failing triangle (-1 if none).
void check_mesh(Ctx& c, const int2* pts, long long nv, const int3* tri, const int3* adj, const int* cflag, int T,
                bool delaunay, int* failed, int* first);
// whole-mesh flip rounds on records already packed (filter_leaves_*_rec);
// the result is unpacked into m.tri / m.adj (and m.cflag when set)
long long flip_rounds_records(Ctx& c, FRec* rec, const Mesh& m, int T, int* rounds);
// spatial order: points sorted by the Morton code of their 1024^2 cell
// (out[k] = pts[perm[k]], inv = perm^-1)
// Counting sort of points by 20-bit Morton cell; cell_end[BITMAP_CELLS + 1]
// receives each cell's end offset in the sorted order.
void morton_order(Ctx& c, const int2* pts, int n, int2* out, int* perm, int* inv, int* cell_end);
// One sample point per occupied 256x256 cell of Morton-ordered points (the
// cell's extreme point towards the nearer edges of the square), from
// morton_order's cell_end; returns the count.
int sample_points_morton(Ctx& c, const int2* pts, const int* cell_end, int* sample_ids);
void relabel_pairs(Ctx& c, const int2* in, int C, const int* inv, int n, int2* out);
void unmap_tris(Ctx& c, int3* tri, int T, const int* perm, int n);
// batches: instance map, per-instance corners/seed triangles/location, local
// constraint ids -> global, and the per-instance output scatter
void batch_vinst(Ctx& c, const long long* off, int K, long long NP, int* vinst);
void batch_init(Ctx& c, int2* pts, int3* tri, int3* adj, int K, long long NP);
// points counting-sorted by instance, then by the Morton cell of a 4^lvl[i]
// grid (cells of instance i numbered from cbase[i]); cell_end[ncells + 1]
// receives each cell's end offset
void batch_order(Ctx& c, const int2* pts, const int* vinst, const int* cbase, const int* lvl, long long NP,
                 long long ncells, int2* out, int* perm, int* inv, int* cell_end);
// the first point of every non-empty 16x16 cell of batch_order's output
int batch_sample_sorted(Ctx& c, const int* cbase, const int* lvl, const int* cell_end, int K, int* ids);
// listed points into their instance's seed triangles
void batch_locate_ids(Ctx& c, const int2* pts, const int* vinst, const int* ids, int n, int* loc, int* pend);
// bitmap of a batch's sample meshes: 16x16 cells per instance; offsets has
// (K*256 + 1) ints
void build_bitmap_batch(Ctx& c, const int2* pts, const int* vinst, long long NP, int K, const int3* tri, int T,
                        int** offsets, int** tris, int* ntris);
int locate_bitmap_batch(Ctx& c, const int2* pts, const int* vinst, long long NP, const int3* tri, int T,
                        const int* bm_off, const int* bm_tri, int* loc, int* isv, int* pend);
void batch_cons(Ctx& c, const int2* in, const long long* coff, const long long* poff, int K, long long NC, int2* out);
// per-instance triangle lists with local vertex ids; perm (nullable) maps the
// batch's internal point order back to the caller's
void batch_out(Ctx& c, const int3* tri, int T, const int* vinst, const int* perm, const long long* poff,
               const long long* toff, long long NP, int K, int3* out);
// indices i with flags[i] != 0 (flags are 0/1), ascending; returns the count
int flags_to_list(Ctx& c, const int* flags, int n, int* list);
// leaves of a DAG -> compact arrays (adjacency remapped); returns count
int filter_leaves(Ctx& c, const int3* tri, const int3* adj, const int* sons, int T, int3* tri2, int3* adj2,
                  long long cap);
// same, with the leaves placed in Morton order of their centroid cell
int filter_leaves_morton(Ctx& c, const int2* pts, const int3* tri, const int3* adj, const int* sons, int T, int3* tri2,
                         int3* adj2, long long cap);
// batch meshes: leaves sorted by (instance, centroid cell in the instance's
// grid of 4^L cells: cbase/lvl/ncells as for the batched point order)
int filter_leaves_inst(Ctx& c, const int2* pts, const int* vinst, const int* cbase, const int* lvl, long long ncells,
                       long long NP, const int3* tri, const int3* adj, const int* sons, int T, int3* tri2, int3* adj2,
                       long long cap);
// the same two orders, written straight into packed flip records
int filter_leaves_morton_rec(Ctx& c, const int2* pts, const int3* tri, const int3* adj, const int* sons, int T,
                             FRec* rec, long long cap);
int filter_leaves_inst_rec(Ctx& c, const int2* pts, const int* vinst, const int* cbase, const int* lvl,
                           long long ncells, long long NP, const int3* tri, const int3* adj, const int* sons, int T,
                           FRec* rec, long long cap);
// bitmap CSR over 1024^2 cells; allocates offsets/tris from the pool
void build_bitmap(Ctx& c, const int2* pts, const int3* tri, int T, int** offsets, int** tris, int* ntris);
// mark vertices of tri[0..T) in isv (must be zeroed), then locate the other
// points of [0,n) through the bitmap; fills loc and pend; returns npend
int locate_bitmap(Ctx& c, const int2* pts, int n, const int3* tri, int T, const int* bm_off, const int* bm_tri,
                  int* loc, int* isv, int* pend);
// locate all points of [0,n) in the two corner triangles 0/1
int locate_two(Ctx& c, const int2* pts, int n, int* loc, int* pend);
int locate_two_ids(Ctx& c, const int2* pts, const int* ids, int n, int* loc, int* pend);
// adjacency of a plain triangle list (edges matched through vertex buckets)
void build_adjacency(Ctx& c, const int3* tri, int T, int nv, int3* adj);
// constrained flags: bit k of cflag[t] set when edge k of t equals a constraint
void mark_constraint_edges(Ctx& c, const int3* tri, const int3* adj, int T, int nv, const int2* cons, int C,
                           int* cflag);
// constraint recovery by cavity retriangulation; touched (optional, T ints)
// receives 1 for every triangle slot that was rewritten
struct ConsResult {
    int rounds = 0;
    long long pieces = 0;
    long long rewritten = 0;
};
// nmax: constraint endpoints must be vertex ids in [0, nmax)
// cell_end (optional): the Morton cell ends of m.pts (morton_order), so the
// split pass finds the vertices on each segment by cells instead of walking
// With `defer`, a piece whose walk meets a constrained edge (it crosses a
// piece recovered before it) is retired instead of failing the solve; the
// pieces and a flag per deferred piece are handed back for the crossing pass.
struct ConsDefer {
    DBuf<int2> pieces;
    DBuf<int> flag;
    int np = 0;
};
ConsResult insert_constraints(Ctx& c, const Mesh& m, int T, int nv, int nmax, const int2* cons, int C, int* touched,
                              const int* cell_end, ConsDefer* defer = nullptr);
// Crossing pass (steiner.hip): inserts the deferred pieces into the CDT of the
// others with a vertex at every crossing (ids nv0, nv0 + 1, ...); the mesh
// arrays hold tcap triangles.  The crossing points are written to out with
// their two pieces mapped to caller ids through perm (ids below n).  Returns
// the new triangle count; *S receives the number of crossing vertices.
int steiner_pass(Ctx& c, const Mesh& m, int T, int tcap, int nv0, const ConsDefer& d, int scap, gcdt_hpoint* out,
                 const int* perm, int n, int* S);
// mark every edge chain of the mesh that runs exactly along a constraint
// (constraints split at collinear vertices) as constrained; returns pieces
long long mark_constraint_chains(Ctx& c, const Mesh& m, int T, int nv, int nmax, const int2* cons, int C);
// vertex ids n..n+3 <- the square corners (insertCornerPoints)
void write_corners(Ctx& c, int2* pts, int n);

// the whole solve (gcdt_solve / gcdt_solve_flags) on device arrays
struct SolveOut {
    int T = 0;
    int S = 0;  // crossing vertices (gcdt_solve_steiner)
    gcdt_stats st{};
};
// crossing constraints resolved (gcdt_solve_steiner): at most scap crossing
// vertices, written to out; the mesh arrays hold 2 (N + scap) + 2 triangles
struct SteinerArgs {
    int scap = 0;
    gcdt_hpoint* out = nullptr;
};
void solve_impl(Ctx& c, const int2* pts_in, int N, const int2* cons_in, int C, int3* tris_out, int3* adj_out,
                SolveOut& out, int* cflag_out = nullptr, const SteinerArgs* steiner = nullptr);
// set-up of gcdt_solve_noalloc on the current device (from gcdt_pool_init)
void noalloc_setup();

}  // namespace gcdt
