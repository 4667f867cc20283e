// Stand-in for <hip/hip_runtime.h> used only by the CPU SIMT emulator
// (tests/emu): the library's HIP sources compile as host C++ with g++, every
// kernel launch becomes a call into the emulator's scheduler, which runs each
// workgroup as fibers (one per work-item) with wave64 shuffles, ballots and
// workgroup barriers, and device allocations are guard-page protected host
// memory.  Test infrastructure only; nothing in the product includes it.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <tuple>
#include <type_traits>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
// __shared__ declarations are rewritten by the emulator build into
// references to per-workgroup storage (emu::lds), so that workgroups that run
// interleaved keep separate LDS
#define __align__(n) alignas(n)
#define __HIP_DEVICE_COMPILE__ 1  // the device branches of the sources are the ones emulated

// ------------------------------------------------------------------ vector types
struct alignas(8) int2 {
    int x, y;
};
struct int3 {
    int x, y, z;
};
struct alignas(16) int4 {
    int x, y, z, w;
};
inline int2 make_int2(int x, int y) { return int2{x, y}; }
inline int3 make_int3(int x, int y, int z) { return int3{x, y, z}; }
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

// ------------------------------------------------------------------ emulator core
namespace emu {
struct U3 {
    unsigned x, y, z;
};
const U3& cur_tid();
const U3& cur_bid();
const U3& cur_bdim();
const U3& cur_gdim();
void block_barrier();
int cur_slot();
void atomic_point();  // a switch point before every atomic (when enabled)
void mem_point();     // a switch point between the parts of a record access (when enabled)
constexpr int kMaxConc = 16;
template <class T, int ID>
inline T& lds()
{
    static T store[kMaxConc];
    return store[cur_slot()];
}
// one wave-wide exchange: op 0 shfl(src), 1 shfl_up(d), 2 shfl_down(d), 3 ballot
unsigned long long wave_op(int op, unsigned long long v, int arg);

struct LaunchCfg {
    dim3 grid, block;
    LaunchCfg(dim3 g, dim3 b, size_t = 0, void* = nullptr) : grid(g), block(b) {}
    template <class S>
    LaunchCfg(dim3 g, dim3 b, size_t sh, S* s) : LaunchCfg(g, b, sh, (void*)s)
    {
    }
};

struct Body {
    virtual void run() = 0;
    virtual ~Body() {}
};
void run_grid(const char* name, const LaunchCfg& cfg, Body& body);

template <class F>
struct BodyOf : Body {
    F f;
    explicit BodyOf(F&& g) : f(std::move(g)) {}
    void run() override { f(); }
};

template <class... P, class... A>
void launch(const char* name, const LaunchCfg& cfg, void (*k)(P...), A&&... a)
{
    std::tuple<std::decay_t<P>...> args(std::forward<A>(a)...);
    auto f = [&]() {
        std::tuple<std::decay_t<P>...> mine = args;  // by-value kernel arguments, per work-item
        std::apply(k, mine);
    };
    BodyOf<decltype(f)> b(std::move(f));
    run_grid(name, cfg, b);
}
}  // namespace emu

#define threadIdx (::emu::cur_tid())
#define blockIdx (::emu::cur_bid())
#define blockDim (::emu::cur_bdim())
#define gridDim (::emu::cur_gdim())
#define GCDT_MEM_POINT() ::emu::mem_point()

inline void __syncthreads() { ::emu::block_barrier(); }
inline int __lane_id() { return (int)(threadIdx.x & 63u); }

template <class T>
inline unsigned long long emu_bits(T v)
{
    static_assert(sizeof(T) <= 8, "wave exchange of at most 8 bytes");
    unsigned long long u = 0;
    memcpy(&u, &v, sizeof(T));
    return u;
}
template <class T>
inline T emu_unbits(unsigned long long u)
{
    T v;
    memcpy(&v, &u, sizeof(T));
    return v;
}
template <class T>
inline T __shfl(T v, int src, int width = 64)
{
    (void)width;
    return emu_unbits<T>(::emu::wave_op(0, emu_bits(v), src));
}
template <class T>
inline T __shfl_up(T v, unsigned d, int width = 64)
{
    (void)width;
    return emu_unbits<T>(::emu::wave_op(1, emu_bits(v), (int)d));
}
template <class T>
inline T __shfl_down(T v, unsigned d, int width = 64)
{
    (void)width;
    return emu_unbits<T>(::emu::wave_op(2, emu_bits(v), (int)d));
}
inline unsigned long long __ballot(int pred) { return ::emu::wave_op(3, pred != 0, 0); }

inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __clzll(long long x) { return x == 0 ? 64 : __builtin_clzll((unsigned long long)x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline unsigned long long __umul64hi(unsigned long long a, unsigned long long b)
{
    return (unsigned long long)(((unsigned __int128)a * b) >> 64);
}

template <class A, class B>
inline std::common_type_t<A, B> min(A a, B b)
{
    return a < b ? a : b;
}
template <class A, class B>
inline std::common_type_t<A, B> max(A a, B b)
{
    return a > b ? a : b;
}

// Atomics: the work-items of a launch run on one host thread, so each atomic
// is an ordinary read-modify-write.
template <class T, class V>
inline T atomicAdd(T* p, V v)
{
    ::emu::atomic_point();
    T o = *p;
    *p = (T)(o + (T)v);
    return o;
}
template <class T, class V>
inline T atomicSub(T* p, V v)
{
    ::emu::atomic_point();
    T o = *p;
    *p = (T)(o - (T)v);
    return o;
}
template <class T, class V>
inline T atomicMax(T* p, V v)
{
    ::emu::atomic_point();
    T o = *p;
    if ((T)v > o) *p = (T)v;
    return o;
}
template <class T, class V>
inline T atomicMin(T* p, V v)
{
    ::emu::atomic_point();
    T o = *p;
    if ((T)v < o) *p = (T)v;
    return o;
}
template <class T, class V>
inline T atomicOr(T* p, V v)
{
    ::emu::atomic_point();
    T o = *p;
    *p = (T)(o | (T)v);
    return o;
}
template <class T, class V>
inline T atomicAnd(T* p, V v)
{
    ::emu::atomic_point();
    T o = *p;
    *p = (T)(o & (T)v);
    return o;
}
template <class T, class V>
inline T atomicExch(T* p, V v)
{
    ::emu::atomic_point();
    T o = *p;
    *p = (T)v;
    return o;
}
template <class T, class U, class V>
inline T atomicCAS(T* p, U cmp, V v)
{
    ::emu::atomic_point();
    T o = *p;
    if (o == (T)cmp) *p = (T)v;
    return o;
}
#define __HIP_MEMORY_SCOPE_SINGLETHREAD 1
#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5
template <class T>
inline T __hip_atomic_load(const T* p, int, int)
{
    ::emu::atomic_point();
    return *p;
}
template <class T, class V>
inline void __hip_atomic_store(T* p, V v, int, int)
{
    *p = (T)v;
}
inline void __threadfence() {}
inline void __threadfence_block() {}

// ------------------------------------------------------------------ runtime API
typedef int hipError_t;
enum {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
};
typedef struct emu_stream_t* hipStream_t;
typedef struct emu_event_t* hipEvent_t;
enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
};
#define hipStreamDefault 0
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0
#define hipHostMallocPortable 1
#define hipHostMallocMapped 2
#define hipHostMallocWriteCombined 4
#define hipHostMallocCoherent 0x40000000
struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
    size_t totalGlobalMem;
};

hipError_t hipMalloc(void** p, size_t bytes);
template <class T>
inline hipError_t hipMalloc(T** p, size_t bytes)
{
    return hipMalloc((void**)p, bytes);
}
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags = 0);
template <class T>
inline hipError_t hipHostMalloc(T** p, size_t bytes, unsigned flags = 0)
{
    return hipHostMalloc((void**)p, bytes, flags);
}
hipError_t hipHostFree(void* p);
hipError_t hipHostGetDevicePointer(void** d, void* h, unsigned flags);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s = nullptr);
hipError_t hipMemset(void* dst, int v, size_t bytes);
hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t s = nullptr);
hipError_t hipStreamCreate(hipStream_t* s);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamQuery(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipGetDevice(int* d);
hipError_t hipSetDevice(int d);
hipError_t hipGetDeviceCount(int* n);
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int d);
hipError_t hipMemGetInfo(size_t* free_bytes, size_t* total_bytes);
// device symbols: a file-scope __device__ array is ordinary host memory here
#define HIP_SYMBOL(x) (&(x))
template <class T>
inline hipError_t hipGetSymbolAddress(void** p, const T* sym)
{
    *p = (void*)sym;
    return hipSuccess;
}
