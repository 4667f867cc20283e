/*
 * gcdt.h -- C ABI of the MI355X (gfx950) constrained Delaunay triangulator.
 *
 * Plain C: pointers and sizes only.  Every array argument is a DEVICE pointer
 * unless the name says host.  Points are on the reference's 2^30 integer grid
 * (reference: include/math_base.cuh:19 logPointRes=30, kernel.cu:171-188) and
 * lie strictly inside (0, 2^30-1)^2; the 4 square corners are vertex ids
 * numPoints..numPoints+3.  Triangles are counter-clockwise vertex triples;
 * adjacency component k is the neighbour across the edge opposite vertex k
 * (-1 none); sons is the insertion DAG (-1 leaf) -- the reference layout of
 * rawMesh.cu:5-22,200-280.
 *
 * Every entry point returns a status (GCDT_OK = 0) instead of the reference's
 * elapsed-time return value; the elapsed GPU milliseconds go to *ms when ms
 * is not NULL.  Buffers the library allocates come from the MemCache pool and
 * are released with gcdt_release(ptr, bytes) (the reference's
 * utils::release(ptr, count), include/utils.cuh:33-39).
 *
 * The (tmp, tmpBytes) pair is the reference's caller-provided CUB scratch
 * (kernel.cu:5386-5387).  Here it is optional: when tmp is NULL the library
 * takes its scan scratch from the pool; when it is given and smaller than
 * gcdt_scratch_bytes(n) the call fails with GCDT_ERR_SCRATCH.
 */
#ifndef GCDT_H
#define GCDT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCDT_OK 0
#define GCDT_ERR_OOM 1            /* device allocation failed */
#define GCDT_ERR_INVALID 2        /* bad argument (range, ids, sizes) */
#define GCDT_ERR_NOT_CONVERGED 3  /* an iterative phase hit its round cap */
#define GCDT_ERR_DEVICE 4         /* HIP runtime error */
#define GCDT_ERR_CROSSING 5       /* two constraint segments cross */
#define GCDT_ERR_SCRATCH 6        /* caller scratch too small */
#define GCDT_ERR_LOCATE 7         /* a point could not be located / duplicate point */
#define GCDT_ERR_CAPACITY 8       /* triangle slot or work-list capacity exceeded */
#define GCDT_ERR_INTERNAL 9       /* mesh consistency check failed */

typedef struct { int32_t x, y; } gcdt_int2;
typedef struct { int32_t x, y, z; } gcdt_int3;

/* ---------------------------------------------------------------- runtime */

/* Last error message of the calling thread (static storage). */
const char* gcdt_last_error(void);
/* Number of visible devices; selects device 0 on first use. */
int gcdt_device_count(int* count);
int gcdt_synchronize(void);
/* "<name>:<gfx arch>:cu<compute units>" of the current device (autotuner cache key). */
int gcdt_device_name(char* buf, int len);

/* MemCache pool (reference: MemCache.cpp:7-30,87-99 getMemByAbtrtSize /
 * releaseAbtrtMem; include/utils.cuh:33-39,65-68 utils::malloc/release). */
int gcdt_alloc(void** ptr, size_t bytes);
int gcdt_release(void* ptr, size_t bytes);
int gcdt_pool_init(void);                 /* MemCache::initAbtrtMem */
int gcdt_pool_stats(size_t* in_use, size_t* cached, size_t* peak);
/* Bytes in use that the pool will not exceed: an allocation past the budget
 * fails with GCDT_ERR_OOM (0 = no budget, the default). */
int gcdt_pool_set_budget(size_t bytes);
/* Restarts the peak that gcdt_pool_stats reports at the bytes now in use. */
int gcdt_pool_reset_peak(void);
/* Returns every cached (free) pool block to the device; blocks in use stay. */
int gcdt_pool_trim(void);
int gcdt_memcpy_h2d(void* dst, const void* host_src, size_t bytes);
int gcdt_memcpy_d2h(void* host_dst, const void* src, size_t bytes);
int gcdt_memcpy_d2d(void* dst, const void* src, size_t bytes);
int gcdt_memset(void* dst, int value, size_t bytes);

/* Non-blocking HIP streams (default priority) for concurrent solves. */
int gcdt_stream_create(void** stream);
int gcdt_stream_destroy(void* stream);
int gcdt_stream_synchronize(void* stream);
int gcdt_memcpy_h2d_async(void* dst, const void* host_src, size_t bytes, void* stream);
int gcdt_memcpy_d2h_async(void* host_dst, const void* src, size_t bytes, void* stream);

/* scratch bytes the (tmp, tmpBytes) argument must offer for n elements */
size_t gcdt_scratch_bytes(int64_t n);

/* Live kernel timing: while enabled, launches of the instrumented kernels
 * (ins_claim, ins_split, ins_fix, ins_reloc, flip_test, flip_apply,
 * flip_fix, locate_bitmap) are bracketed by hipEvents on their stream.
 * gcdt_prof_read returns the summed milliseconds, launch count and items
 * processed (list entries) of one kernel; gcdt_prof_names writes the
 * ';'-separated names seen so far and returns the full length. */
int gcdt_prof_enable(int on);
int gcdt_prof_reset(void);
/* Time only the named kernel (NULL or "": all): fewer event markers in the
 * stream when one kernel's live duration is wanted inside a timed region. */
int gcdt_prof_only(const char* kernel);
int gcdt_prof_read(const char* kernel, double* total_ms, int64_t* launches, int64_t* items);
int gcdt_prof_names(char* buf, int cap);

/* ------------------------------------------------- reference entry points */

/* rawMesh.h:6, rawMesh.cu:713-746.  One point per occupied 256x256 cell of
 * the 2^30 grid (the smallest point id in the cell).  Allocates
 * *samplePoints (gcdt_int2[65536+4]) and *sample (int[65536+4]) from the pool. */
int gcdt_get_sample_points(void* tmp, size_t tmpBytes, const gcdt_int2* points, int numPoints,
                           gcdt_int2** samplePoints, int** sample, int* num, double* ms);

/* rawMesh.h:4, rawMesh.cu:1040-1113.  Writes the 4 corners to
 * points[numPoints..+3] and inserts all points by parallel 1->3 / 1->2
 * splits of the 2 corner triangles.  tris/sons capacity 4*numPoints+4. */
int gcdt_raw_mesh_before(void* tmp, size_t tmpBytes, gcdt_int2* points, gcdt_int3* tris, int* sons,
                         int* numTriangles, int numPoints, double* ms);

/* rawMesh.h:5, rawMesh.cu:994-1038.  tris[0..*numTriangles) with adjTris is
 * a triangulation of some of the points (the converted sample mesh); the
 * remaining points are located with the bitmap and inserted.  Leaves
 * (sons == -1) end with exact adjacency.  Capacity 4*numPoints+4 slots. */
int gcdt_raw_mesh_after(void* tmp, size_t tmpBytes, gcdt_int2* points, gcdt_int3* tris,
                        gcdt_int3* adjTris, int* sons, const int* bitmapOffsets, const int* bitmapTriIds,
                        int* numTriangles, int numPoints, double* ms);

/* rawMesh.h:7, rawMesh.cu:1114-1189.  rawMeshBefore with adjacency. */
int gcdt_raw_mesh(void* tmp, size_t tmpBytes, gcdt_int2* points, gcdt_int3* tris, gcdt_int3* adjTris,
                  int* sons, int* numTriangles, int numPoints, double* ms);

/* boostmap.h:3, boostmap.cu:146-181.  CSR over the 1024x1024 cells of the
 * 2^30 grid: cell (x>>20)*1024+(y>>20) lists every triangle whose closed
 * region meets the closed cell square.  Allocates *bitmapOffsets
 * (int[2^20+1]) and *bitmapTris (int[(*bitmapOffsets)[2^20]]). */
int gcdt_get_bitmap(void* tmp, size_t tmpBytes, const gcdt_int2* points, const gcdt_int3* tris,
                    int** bitmapOffsets, int** bitmapTris, int numTriangles, double* ms);

/* kernel.cu:5293-5324.  Compacts leaves (sons == -1) of tris into newTris
 * (and adjTris into newAdjTris with ids remapped, when both are non-NULL). */
int gcdt_filter_triangles(void* tmp, size_t tmpBytes, const gcdt_int3* tris, gcdt_int3* newTris,
                          const gcdt_int3* adjTris, gcdt_int3* newAdjTris, const int* sons,
                          int* numTriangles, double* ms);

/* kernel.cu:833-860.  Sample-local vertex ids -> global ids; the sample's
 * corners numSamples..+3 -> numPoints..+3. */
int gcdt_convert_sample(gcdt_int3* tris, const int* sampleId, int numSamples, int numPoints, int numTriangles);

/* kernel.cu:5038-5107.  Flips a triangulation of points[0..numPoints) to
 * the Delaunay triangulation in place and writes its adjacency. */
int gcdt_delaunay_sample(void* tmp, size_t tmpBytes, const gcdt_int2* points, gcdt_int3* tris,
                         gcdt_int3* adjTris, int numPoints, int numTriangles, double* ms);

/* fastDelaunay.h:5, fastDelaunay.cu:710-863 (a flip-only host prototype in
 * the reference that never writes back).  Here: flips the given mesh to
 * Delaunay on the GPU, in place, updating adjTris. */
int gcdt_try_fast_delaunay(void* tmp, size_t tmpBytes, const gcdt_int2* points, gcdt_int3* tris,
                           gcdt_int3* adjTris, int numPoints, int numTriangles, double* ms);

/* kernel.cu:4697-4860.  Constrained Delaunay flipping of tris (no adjacency
 * given): builds adjacency, marks every triangle edge that equals a
 * constraint as unflippable, and flips all other edges to local Delaunay.
 * Unlike the reference it does not Morton-reorder points, so *points and
 * *tris keep their addresses and vertex ids. */
int gcdt_delaunay_new(void* tmp, size_t tmpBytes, gcdt_int2** points, gcdt_int3** tris,
                      const gcdt_int2* cons, int numPoints, int numTriangles, int numCons, double* ms);

/* addConstraints.h:5, addConstraints.cu:1723-1835.  Recovers every
 * constraint (vertex-id pair) in the leaf mesh of tris/adjTris/sons.  A
 * segment that passes exactly through vertices is split at them (the
 * reference drops it, addConstraints.cu:962-965).  Triangles crossed by a
 * segment are retired (sons = 0) and their replacement triangles are returned
 * in *newTris (allocated from the pool, *numNewTris entries), exactly as the
 * reference's caller expects (kernel.cu:5427-5442). */
int gcdt_add_constraints(void* tmp, size_t tmpBytes, const gcdt_int2* points, gcdt_int3* tris,
                         gcdt_int3* adjTris, int* sons, const int* bitmapOffsets, const int* bitmapTris,
                         const gcdt_int2* cons, gcdt_int3** newTris, int numCons, int numPoints,
                         int numTriangles, int* numNewTris, double* ms);

/* -------------------------------------------------------------- one call */

typedef struct {
    double ms_total;       /* GPU time of the whole solve */
    double ms_sample;      /* sampling + sample mesh + sample flips */
    double ms_bitmap;
    double ms_insert;      /* location + insertion rounds */
    double ms_flip;        /* Delaunay flips of the full mesh */
    double ms_constraints; /* constraint recovery + constrained flips */
    int insert_rounds;
    int flip_rounds;
    int constraint_rounds;
    int num_triangles;
    int64_t flips;
    int64_t segments;      /* constraint pieces after splitting at vertices */
    int64_t splits;        /* pieces - constraints */
} gcdt_stats;

/* Full CDT of points[0..N) (device, 2^30 grid, strictly inside the square,
 * pairwise distinct) with constraints cons[0..C) (device vertex-id pairs,
 * non-crossing).  Writes 2N+2 triangles to tris_out (device, capacity 2N+2),
 * vertex ids as given, corners N..N+3.  adj_out (optional, may be NULL)
 * receives the adjacency.  stream may be NULL (default stream). */
int gcdt_solve(const gcdt_int2* points, int N, const gcdt_int2* cons, int C, gcdt_int3* tris_out,
               gcdt_int3* adj_out, int* num_tris, gcdt_stats* stats, void* stream);

/* gcdt_solve that also writes the constraint bits of every output triangle
 * (bit k of cflag_out[t] set: the edge opposite vertex k lies on a
 * constraint); adj_out and cflag_out are required. */
int gcdt_solve_flags(const gcdt_int2* points, int N, const gcdt_int2* cons, int C, gcdt_int3* tris_out,
                     gcdt_int3* adj_out, int* cflag_out, int* num_tris, gcdt_stats* stats, void* stream);

/* gcdt_solve that allocates nothing: no counterpart in the reference.  It is
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Every buffer
 * of the solve comes from a static device array of GCDT_NOALLOC_SCRATCH_BYTES
 * per device (`scratch`, gcdt_amd/csrc/noalloc.hip), held from library load
 * until the process ends; gcdt_pool_init() must have run once on the device
 * first (it looks up the array's device address and reserves pinned counter
 * pages), else GCDT_ERR_INVALID.  Arguments, outputs and status codes are
 * those of gcdt_solve.  The array holds about 40 MB plus 260 bytes per point,
 * plus about 48 bytes per triangle that a constraint segment crosses: inputs
 * above GCDT_NOALLOC_MAX_POINTS points or segments fail with
 * GCDT_ERR_SCRATCH at once, and an input whose constraints need more than the
 * array fails with GCDT_ERR_SCRATCH naming the bytes it needed.  All calls on
 * a device share its array: they are serialised, each returning after its
 * stream has drained.  stats->ms_* come from the host clock. */
#define GCDT_NOALLOC_SCRATCH_BYTES (256ull << 20)
#define GCDT_NOALLOC_MAX_POINTS 750000
int gcdt_solve_noalloc(const gcdt_int2* points, int N, const gcdt_int2* cons, int C, gcdt_int3* tris_out,
                       gcdt_int3* adj_out, int* num_tris, gcdt_stats* stats, void* stream);

/* A crossing point of two constraint segments, exactly: (X / w, Y / w) with
 * X = x_hi * 2^64 + x_lo and Y likewise (128-bit two's complement), w > 0.
 * On the 2^30 grid |X|, |Y| < 2^92 and w < 2^61.  seg holds the vertex ids of
 * two constraint pieces through it, (seg[0], seg[1]) and (seg[2], seg[3]). */
typedef struct {
    int64_t x_hi;
    uint64_t x_lo;
    int64_t y_hi;
    uint64_t y_lo;
    int64_t w;
    int32_t seg[4];
} gcdt_hpoint;

/* gcdt_solve_flags with crossing constraints resolved instead of refused
 * (GCDT_ERR_CROSSING): no counterpart in the reference, whose walk drops a
 * segment that meets a constrained edge (addConstraints.cu:768-774,942-946).
 * Every proper crossing of two constraint pieces becomes a vertex (ids N+4,
 * N+5, ...; coordinates in steiner_out[0..*num_steiner)) and each piece a
 * chain of constrained edges through its crossings; the result is the CDT of
 * the points, the corners and the crossings.  tris_out, adj_out and cflag_out
 * hold 2 (N + max_steiner) + 2 entries (adj_out and cflag_out may be NULL);
 * *num_tris = 2 (N + *num_steiner) + 2.  More than max_steiner crossings fail
 * with GCDT_ERR_CAPACITY.  The crossings are inserted by one device thread,
 * one piece after another: meant for inputs with few crossings. */
int gcdt_solve_steiner(const gcdt_int2* points, int N, const gcdt_int2* cons, int C, int max_steiner,
                       gcdt_int3* tris_out, gcdt_int3* adj_out, int* cflag_out, gcdt_hpoint* steiner_out,
                       int* num_tris, int* num_steiner, gcdt_stats* stats, void* stream);

/* Even-odd carving (no counterpart in the reference, which keeps every
 * triangle of the bounding square): inside[t] = 1 for the triangles enclosed
 * by an odd number of constraint boundaries -- polygon interiors minus their
 * holes -- counted from the region holding the square's corners (vertex ids
 * >= numPoints).  Open polylines separate nothing and change no parity.
 * tris/adj/cflag as written by gcdt_solve_flags; all arrays on the device. */
int gcdt_carve(const gcdt_int3* tris, const gcdt_int3* adj, const int* cflag, int T, int numPoints, int* inside,
               int* num_inside, void* stream);

/* Mesh check on the device (no counterpart in the reference, which only
 * prints debug counts; the solves run it after every phase when the
 * "validate" tunable is 1): points holds nv vertices (corners included),
 * tris/adj T triangles with -1 for the bounding square's sides, cflag (or
 * NULL) the constraint bits of gcdt_solve_flags.  *failed gets a bit set per
 * failed check -- 1 vertex ids, 2 orientation (not counter-clockwise),
 * 4 adjacency (a neighbour that does not name the triangle back across the
 * same edge), 8 constraint bits differing across an edge, 16 an unconstrained
 * edge not locally Delaunay (checked when delaunay is 1) -- and *first_bad the
 * lowest failing triangle (-1 if none).  Returns GCDT_OK whatever it finds. */
int gcdt_check_mesh(const gcdt_int2* points, long long nv, const gcdt_int3* tris, const gcdt_int3* adj,
                    const int* cflag, int T, int delaunay, int* failed, int* first_bad, void* stream);

/* ------------------------------------------------------------ batches */

typedef struct {
    double ms_total;
    double ms_insert;      /* location + insertion rounds of all instances */
    double ms_flip;
    double ms_constraints;
    int insert_rounds;
    int flip_rounds;
    int constraint_rounds;
    int num_instances;
    int64_t num_points;
    int64_t num_triangles;
    int64_t flips;
    int64_t segments;
} gcdt_batch_stats;

/* K independent CDTs in one set of launches.  Instance i owns points
 * [point_offsets[i], point_offsets[i+1]) of `points` (device, 2^30 grid,
 * distinct within the instance) and constraints [cons_offsets[i],
 * cons_offsets[i+1]) of `cons` (device, vertex ids LOCAL to the instance).
 * The offsets arrays are HOST arrays of K+1 entries (cons_offsets may be
 * NULL when there are no constraints).  Instance i's 2*N_i+2 triangles go to
 * tris_out[sum_{j<i} (2*N_j+2) ...] with local vertex ids (corners
 * N_i..N_i+3).  Every round of every phase runs for all instances at once,
 * so a batch costs about as many rounds as its largest instance.  Like
 * gcdt_solve, each instance first triangulates a sample (one point per cell
 * of its own 16x16 grid), makes it Delaunay and locates the other points
 * through a per-instance bitmap of it.  K <= 262144.
 * inst_status (optional, HOST array of K ints): with it, an instance whose
 * own input makes the solve fail (bad ids or coordinates, duplicate points,
 * crossing constraints) fails alone: the batch is split in halves down to
 * that instance, inst_status[i] receives each instance's GCDT_* status, the
 * other instances' triangles are written as usual, the call returns GCDT_OK
 * and gcdt_last_error() names the first failed instance.  Without it, or
 * for device and memory errors, the whole call fails. */
int gcdt_solve_batch(const gcdt_int2* points, const int64_t* point_offsets, int K, const gcdt_int2* cons,
                     const int64_t* cons_offsets, gcdt_int3* tris_out, gcdt_batch_stats* stats, int* inst_status,
                     void* stream);

/* Process-wide tunables read by every later solve (set them between solves,
 * not while a solve runs on another thread): "tail_threshold" [0, 2^20] (list
 * length at or below which rounds finish in one persistent workgroup),
 * "grid_cap" [256, 65536] (workgroups of a grid-stride launch),
 * "flip_items" / "ins_items" (1, 2 or 4 list items per thread of the flip
 * test/apply and insertion split/relocate launches), "pivot"
 * (0 random, 1 central-first, the default), "single_tester" (0/1), "validate"
 * (0/1: 1 checks the mesh on the device after every phase -- orientation,
 * adjacency symmetry, constraint bits on both sides of an edge, local Delaunay
 * after the flip phases -- and fails the solve with GCDT_ERR_INTERNAL naming
 * the phase; environment GCDT_VALIDATE).  Out-of-range values
 * fail with GCDT_ERR_INVALID, and so do in-range values that no GPU parity
 * test covers (the same rule applies to the GCDT_* environment overrides: a
 * refused one makes every solve fail).  gcdt_amd/tune.py picks tail_threshold,
 * grid_cap and the items per thread by timing and caches the choice per
 * device. */
int gcdt_set_tunable(const char* name, int value);
int gcdt_get_tunable(const char* name, int* value);

/* ------------------------------------------------------------ I/O formats */

/* kernel.cu:5159-5291 readFile.  Text file: a point count, then one line per
 * point "idx x y" (idx in [0,count), x and y doubles), then an optional
 * constraint count and one line per constraint "k a b" (a, b point idx).
 * Points with identical coordinates are merged (the first keeps its slot,
 * later idx map onto it); the rest are scaled about the bounding-box centre
 * by the longer side into [0.005, 0.995]^2 as ((v-c)/l)*0.99+0.5.  With
 * shuffle != 0 point i moves to slot random[i], random = std::shuffle of
 * 0..n-1 under std::mt19937(19999) (the permutation is the host C++ library's).
 * Constraints are mapped through the merge (and the shuffle); self-loops and
 * repeats in either orientation are dropped.  Outputs are HOST arrays
 * (2*numPoints doubles, numCons pairs) released with gcdt_free_host.
 * Deviations: the bounding box starts from the first point (the reference
 * starts from +-1e6, which is wrong when every coordinate lies beyond 1e6), a
 * zero extent scales by 1, and malformed files fail with GCDT_ERR_INVALID. */
int gcdt_read_file(const char* path, int shuffle, double** points, int* numPoints, gcdt_int2** cons,
                   int* numCons);
void gcdt_free_host(void* ptr);

/* kernel.cu:171-188 transPoints: device double[2n] -> device int2[n] on the
 * grid of resolution res (2^30 in the reference): trunc(v*res) clamped to
 * [0, res-1]. */
int gcdt_trans_points(const double* fpoints, int n, long long res, gcdt_int2* points, void* stream);

/* Canonical triangle list, in place on HOST memory: each triangle rotated so
 * that its smallest vertex id comes first (orientation kept), then sorted
 * lexicographically.  Two triangulations of the same points are equal iff
 * their canonical lists are. */
int gcdt_canonicalize_host(gcdt_int3* tris, int64_t numTriangles);

/* Writes HOST points and triangles as text: the point count, "i x y" lines,
 * the triangle count, "i a b c" lines (the readFile layout, extended). */
int gcdt_write_mesh(const char* path, const gcdt_int2* points, int numPoints, const gcdt_int3* tris,
                    int64_t numTriangles);

#ifdef __cplusplus
}
#endif

#endif /* GCDT_H */
