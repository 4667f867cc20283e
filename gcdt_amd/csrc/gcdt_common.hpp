// Shared definitions for the gfx950 CDT kernels: status codes, error checks,
// launch geometry and wave64 helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/gcdt.h"

// A point between the separate memory instructions of one multi-word record
// load or store.  It generates nothing: on the GPU another work-item's
// accesses can land between those instructions anyway (a record read
// concurrently with its rewrite may see some 16-byte parts old and others
// new).  The CPU SIMT emulator of the test suite defines it as a switch
// between work-items, so that such interleavings happen on the host too.
#ifndef GCDT_MEM_POINT
#define GCDT_MEM_POINT()
#endif

namespace gcdt {

// hipError_t -> exception carrying a GCDT status; caught at the C-ABI edge.
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define GCDT_HIP(expr)                                                                   \
    do {                                                                                 \
        hipError_t e__ = (expr);                                                         \
        if (e__ != hipSuccess) {                                                         \
            throw ::gcdt::Error(e__ == hipErrorOutOfMemory ? GCDT_ERR_OOM : GCDT_ERR_DEVICE, \
                                std::string(#expr) + ": " + hipGetErrorString(e__));     \
        }                                                                                \
    } while (0)

#define GCDT_LAUNCH_CHECK() GCDT_HIP(hipGetLastError())

constexpr int kBlock = 256;         // 4 wave64 per workgroup
constexpr int kMaxGrid = 256 * 16;  // persistent grid-stride cap: 16 WGs per CU

// Cap on workgroups of a grid-stride launch (a tunable; kMaxGrid by default).
extern int g_grid_cap;

inline int grid_for(long long n)
{
    long long g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > g_grid_cap) g = g_grid_cap;
    return (int)g;
}

// Wave-uniform grid-stride loop over [0, n): every lane of a wave stays in the
// loop for the same iterations (lanes past the end see valid == false), so
// wave-level ballots and shuffles inside the body are always converged.
#define WAVE_LOOP(i, n)                                                                    \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x,                   \
                   i##_wb = i - (long long)__lane_id();                                    \
         i##_wb < (long long)(n);                                                          \
         i += (long long)gridDim.x * blockDim.x, i##_wb += (long long)gridDim.x * blockDim.x)

// Exclusive prefix sum of v across the 64 lanes; all lanes must be active.
__device__ inline int wave_excl_scan(int v, int& total)
{
    int lane = __lane_id();
    int incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    total = __shfl(incl, 63, 64);
    return incl - v;
}

// Reserve k consecutive entries of a device-wide counter per lane with one
// atomic per wave; all lanes must be active (k = 0 for lanes that need none).
__device__ inline int wave_alloc(int k, int* counter)
{
    int total;
    int ex = wave_excl_scan(k, total);
    int base = 0;
    if (__lane_id() == 0 && total > 0) base = atomicAdd(counter, total);
    base = __shfl(base, 0, 64);
    return base + ex;
}

// atomicAdd(&ctr[key], 1) for every lane with key >= 0, combined over runs of
// equal keys in consecutive lanes: the first lane of a run adds the run's
// length once and each lane gets the value its own add would have returned in
// lane order.  Keys < 0 add nothing and get -1.  All lanes must be active.
__device__ inline int wave_run_add(int* ctr, int key)
{
    int lane = __lane_id();
    int prev = __shfl_up(key, 1, 64);
    bool head = lane == 0 || prev != key;
    unsigned long long heads = __ballot(head);
    unsigned long long upto = (2ull << lane) - 1;  // lanes 0..lane (all 64 for lane 63)
    int first = 63 - __clzll(heads & upto);
    unsigned long long after = heads & ~upto;
    int next = after ? __ffsll((long long)after) - 1 : 64;
    int base = 0;
    if (head && key >= 0) base = atomicAdd(&ctr[key], next - lane);
    base = __shfl(base, first, 64);
    return key >= 0 ? base + lane - first : -1;
}

// Block-wide reservation of k entries per thread with one atomicAdd per block.
// Every thread of the block must call it (block-uniform control flow).
template <int BS>
__device__ inline int block_alloc(int k, int* counter)
{
    __shared__ int s_w[BS / 64];
    __shared__ int s_base;
    int lane = __lane_id(), w = threadIdx.x >> 6;
    int wt;
    int ex = wave_excl_scan(k, wt);
    if (lane == 0) s_w[w] = wt;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < BS / 64; i++) {
        int v = s_w[i];
        if (i < w) pre += v;
        tot += v;
    }
    if (threadIdx.x == 0) s_base = tot > 0 ? atomicAdd(counter, tot) : 0;
    __syncthreads();
    int base = s_base;
    __syncthreads();
    return base + pre + ex;
}

// Block-wide exclusive prefix of k; *tot receives the block's total (the same
// in every thread).  Every thread of the block must call it.
template <int BS>
__device__ inline int block_excl_prefix(int k, int* tot)
{
    __shared__ int s_w[BS / 64];
    int lane = __lane_id(), w = threadIdx.x >> 6;
    int wt;
    int ex = wave_excl_scan(k, wt);
    if (lane == 0) s_w[w] = wt;
    __syncthreads();
    int pre = 0, t = 0;
#pragma unroll
    for (int i = 0; i < BS / 64; i++) {
        int v = s_w[i];
        if (i < w) pre += v;
        t += v;
    }
    __syncthreads();  // s_w is reused by the next call
    *tot = t;
    return pre + ex;
}

// A block's appends to a device list, staged in LDS.  Device atomics on one
// address serialise (about 11 ns apart on MI355X whatever the issuing
// pattern), so one append atomic per 256-item tile caps a launch near 23 G
// items/s and makes the list counter the critical path of the round kernels.
// Entries of successive tiles collect in buf[CAP] and leave in one contiguous,
// coalesced chunk per flush, with one atomic per chunk.  Entries past `cap`
// are dropped and flag MERR-style `capbit` in *err.  Block-uniform: every
// thread calls reserve() and flush() the same number of times; `staged` holds
// the same value in every thread.
template <int CAP>
struct Staged {
    int* buf;   // LDS, CAP ints
    int* base;  // LDS, 1 int
    int* out;
    int* counter;
    long long cap;
    int* err;
    int capbit;
    int staged;

    __device__ Staged(int* b, int* bs, int* o, int* ctr, long long c = (1LL << 40), int* e = nullptr, int cb = 0)
        : buf(b), base(bs), out(o), counter(ctr), cap(c), err(e), capbit(cb), staged(0)
    {
    }

    __device__ void flush()
    {
        __syncthreads();  // the staged entries are written
        if (staged == 0) return;
        if (threadIdx.x == 0) *base = atomicAdd(counter, staged);
        __syncthreads();
        long long b = *base;
        if (b + staged > cap && threadIdx.x == 0 && err) atomicOr(err, capbit);
        for (int i = threadIdx.x; i < staged; i += blockDim.x)
            if (b + i < cap) out[b + i] = buf[i];
        __syncthreads();  // buf is free again
        staged = 0;
    }

    // This thread's first position in buf for its k entries (k * blockDim.x
    // must not exceed CAP).
    __device__ int reserve(int k)
    {
        int tot;
        int pre = block_excl_prefix<kBlock>(k, &tot);
        if (staged + tot > CAP) flush();
        int p = staged + pre;
        staged += tot;
        return p;
    }
};

// LDS staging capacity of a list append (ints per list per block): at least
// the most entries one tile can append (256 threads x 4 items x 2 entries).
constexpr int STAGE_CAP = 2048;

// Block-uniform tile loop: each iteration covers `per` consecutive items.
#define TILE_LOOP(base, n, per) \
    for (long long base = (long long)blockIdx.x * (per); base < (long long)(n); base += (long long)gridDim.x * (per))

__device__ __host__ inline int nx3(int k) { return k == 2 ? 0 : k + 1; }
__device__ __host__ inline int pv3(int k) { return k == 0 ? 2 : k - 1; }

// by value: with a reference, the selects fold into a computed field offset
// and the caller's local int3 lands on the stack
__device__ __host__ inline int get3(int3 t, int k) { return k == 0 ? t.x : (k == 1 ? t.y : t.z); }
__device__ __host__ inline void set3(int3& t, int k, int v)
{
    if (k == 0) t.x = v;
    else if (k == 1) t.y = v;
    else t.z = v;
}
// slot of vertex v in t, or -1
__device__ __host__ inline int slot_of(const int3& t, int v) { return t.x == v ? 0 : (t.y == v ? 1 : (t.z == v ? 2 : -1)); }
// slot of the vertex of t that is neither u nor v
__device__ __host__ inline int slot_other(const int3& t, int u, int v)
{
    if (t.x != u && t.x != v) return 0;
    if (t.y != u && t.y != v) return 1;
    return 2;
}
__device__ __host__ inline bool has2(const int3& t, int u, int v)
{
    return (t.x == u || t.y == u || t.z == u) && (t.x == v || t.y == v || t.z == v);
}

// Adjacency entries written by a remeshing step before the outer neighbour is
// known to be final: -(old_neighbour + 2).  -1 stays "no neighbour".
__device__ __host__ inline int adj_mark(int m) { return m < 0 ? -1 : -(m + 2); }
__device__ __host__ inline bool adj_is_mark(int a) { return a <= -2; }
__device__ __host__ inline int adj_unmark(int a) { return -a - 2; }

__device__ __host__ inline uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

}  // namespace gcdt
