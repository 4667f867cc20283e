// CPU SIMT emulator for the library's kernels (test infrastructure).
//
// Every launch runs on the calling host thread.  Workgroups run one after
// another (in a seeded shuffled order); the work-items of a workgroup are
// fibers with their own stacks, switched at every wave-wide operation
// (shuffle, ballot) and workgroup barrier, and resumed in a seeded shuffled
// order between those points.  A wave operation completes for the lanes of
// the wave that reached it once no work-item of the workgroup can run on
// (lanes waiting elsewhere are the inactive lanes of that operation); a
// barrier completes once every live work-item reached it.
//
// Device memory is host memory bracketed by inaccessible guard pages: an
// allocation sits against its upper guard (EMU_ALIGN_LOW=0, the default) or
// its lower one (EMU_ALIGN_LOW=1), so an access one element past either end
// faults.  A fault prints the kernel, workgroup and work-item that made it
// and the allocation nearest the address, then exits with status 86.  New
// allocations are filled with seeded random bytes, as a reused pool block
// holds stale values.  Freed memory is unmapped, so a later use faults too.
// The pool (MemCache) hands out exact-size allocations and checks that each
// release names a size of the same size class as the request.
#include <hip/hip_runtime.h>
#include <signal.h>
#include <sys/mman.h>
#include <unistd.h>

#include <chrono>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/MemCache.h"
#include "mesh.hpp"  // the rewritten copy in _build/src

namespace emu {

// ------------------------------------------------------------------ settings

static unsigned long long g_seed = 1;
static int g_align_low = 0;

static unsigned long long splitmix(unsigned long long& s)
{
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// ------------------------------------------------------------------ memory

struct Alloc {
    char* base;       // mapping start (lower guard page)
    size_t map_bytes;
    char* user;
    size_t bytes;
    unsigned long long id;
};
static std::mutex g_mem_mu;
static std::map<uintptr_t, Alloc> g_allocs;  // by user address
static unsigned long long g_alloc_id = 0;
static unsigned long long g_fill_state = 12345;
static const size_t kPage = 4096;

static void* mem_alloc(size_t bytes, bool fill)
{
    if (bytes == 0) bytes = 1;
    size_t data_pages = (bytes + kPage - 1) / kPage;
    size_t map_bytes = (data_pages + 2) * kPage;
    char* base = (char*)mmap(nullptr, map_bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE,
                             -1, 0);
    if (base == MAP_FAILED) return nullptr;
    mprotect(base, kPage, PROT_NONE);
    mprotect(base + (data_pages + 1) * kPage, kPage, PROT_NONE);
    char* user;
    if (g_align_low) user = base + kPage;
    else user = base + (data_pages + 1) * kPage - bytes;
    if (fill) {
        unsigned long long s = splitmix(g_fill_state) ^ g_seed;
        size_t i = 0;
        for (; i + 8 <= bytes; i += 8) {
            unsigned long long r = splitmix(s);
            memcpy(user + i, &r, 8);
        }
        for (; i < bytes; i++) user[i] = (char)splitmix(s);
    }
    std::lock_guard<std::mutex> g(g_mem_mu);
    g_allocs[(uintptr_t)user] = Alloc{base, map_bytes, user, bytes, ++g_alloc_id};
    return user;
}

static size_t mem_free(void* p)
{
    if (!p) return 0;
    std::lock_guard<std::mutex> g(g_mem_mu);
    auto it = g_allocs.find((uintptr_t)p);
    if (it == g_allocs.end()) {
        fprintf(stderr, "emu: free of %p, which is not an allocation\n", p);
        fflush(stderr);
        _exit(87);
    }
    size_t bytes = it->second.bytes;
    munmap(it->second.base, it->second.map_bytes);
    g_allocs.erase(it);
    return bytes;
}

static size_t mem_size(const void* p)
{
    std::lock_guard<std::mutex> g(g_mem_mu);
    auto it = g_allocs.find((uintptr_t)p);
    return it == g_allocs.end() ? 0 : it->second.bytes;
}

// ------------------------------------------------------------------ fibers

extern "C" void emu_ctx_switch(void** save_sp, void* load_sp);
asm(R"(
    .text
    .globl emu_ctx_switch
    .type emu_ctx_switch,@function
emu_ctx_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_ctx_switch, .-emu_ctx_switch
)");

enum State { READY, AT_WAVE, AT_BARRIER, DONE };

struct Fiber {
    void* sp = nullptr;
    char* stack = nullptr;
    size_t stack_bytes = 0;
    U3 tid{0, 0, 0};
    int slot = 0;  // which of the concurrently running workgroups
    State st = DONE;
    int op = 0;
    unsigned long long val = 0, res = 0;
    int arg = 0;
    int stall = 0;  // scheduler passes to skip this work-item (EMU_YIELD_MEM=2)
};

struct Slot {
    U3 bid{0, 0, 0};
    bool active = false;
};

struct Exec {
    const char* name = "";
    Body* body = nullptr;
    U3 bdim{1, 1, 1}, gdim{1, 1, 1};
    int nthreads = 1;
    std::vector<Fiber> fibers;  // slot s holds fibers [s*nthreads, (s+1)*nthreads)
    std::vector<Slot> slots;
    Fiber* cur = nullptr;
    void* sched_sp = nullptr;
};

static thread_local Exec* t_exec = nullptr;
static const U3 kZero{0, 0, 0}, kOne{1, 1, 1};
static std::recursive_mutex g_launch_mu;  // launches from several host threads take turns
static int g_conc = 1;                    // workgroups interleaved at once
static int g_yield_atomics = 0;           // switch work-items at every atomic
static thread_local unsigned long long t_order_state = 1;  // scheduling order
static int g_yield_mem = 0;               // ... and between the 16-byte parts of a record load or store

const U3& cur_tid() { return t_exec && t_exec->cur ? t_exec->cur->tid : kZero; }
const U3& cur_bid() { return t_exec && t_exec->cur ? t_exec->slots[t_exec->cur->slot].bid : kZero; }
const U3& cur_bdim() { return t_exec ? t_exec->bdim : kOne; }
const U3& cur_gdim() { return t_exec ? t_exec->gdim : kOne; }
int cur_slot() { return t_exec && t_exec->cur ? t_exec->cur->slot : 0; }

static void yield_to_scheduler()
{
    Exec* x = t_exec;
    emu_ctx_switch(&x->cur->sp, x->sched_sp);
}

void atomic_point()
{
    Exec* x = t_exec;
    if (!g_yield_atomics || !x || !x->cur) return;
    yield_to_scheduler();  // stays READY: another work-item may run first
}

void mem_point()
{
    Exec* x = t_exec;
    if (!g_yield_mem || !x || !x->cur) return;
    // mode 2: now and then the work-item also stalls for up to 64 passes, so
    // that others run whole bodies between two parts of its record access
    if (g_yield_mem >= 2 && splitmix(t_order_state) % 8 == 0) x->cur->stall = 1 + (int)(splitmix(t_order_state) % 64);
    yield_to_scheduler();
}

[[noreturn]] static void fiber_main()
{
    Exec* x = t_exec;
    x->body->run();
    x->cur->st = DONE;
    emu_ctx_switch(&x->cur->sp, x->sched_sp);
    __builtin_unreachable();
}

static const size_t kStack = 512 * 1024;

static void fiber_prepare(Fiber& f)
{
    if (!f.stack) {
        f.stack_bytes = kStack;
        f.stack = (char*)mmap(nullptr, kStack + kPage, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE,
                              -1, 0);
        if (f.stack == MAP_FAILED) {
            fprintf(stderr, "emu: fiber stack allocation failed\n");
            _exit(88);
        }
        mprotect(f.stack, kPage, PROT_NONE);  // stack overflow faults
    }
    // initial frame: six callee-saved registers, then the return address
    // fiber_main is entered through (rsp = 8 mod 16 at its first instruction)
    uintptr_t top = (uintptr_t)(f.stack + kPage + f.stack_bytes);
    top &= ~(uintptr_t)15;
    void** sp = (void**)top;
    *--sp = nullptr;  // alignment slot
    *--sp = (void*)&fiber_main;
    for (int i = 0; i < 6; i++) *--sp = nullptr;
    f.sp = sp;
    f.st = READY;
    f.stall = 0;
}

void block_barrier()
{
    Exec* x = t_exec;
    if (!x || !x->cur) return;
    x->cur->st = AT_BARRIER;
    yield_to_scheduler();
}

unsigned long long wave_op(int op, unsigned long long v, int arg)
{
    Exec* x = t_exec;
    if (!x || !x->cur) {
        // host-side call (no launch): a one-lane wave
        return op == 3 ? (v ? 1ull : 0ull) : v;
    }
    Fiber* f = x->cur;
    f->st = AT_WAVE;
    f->op = op;
    f->val = v;
    f->arg = arg;
    yield_to_scheduler();
    return f->res;
}

static void die(const char* what)
{
    Exec* x = t_exec;
    fprintf(stderr, "emu: %s in kernel %s\n", what, x ? x->name : "?");
    fflush(stderr);
    _exit(89);
}

// Completes the wave operation of wave w of slot s once none of its lanes
// can run on; returns whether it did.
static bool release_wave(Exec* x, int s, int w)
{
    const int nthreads = x->nthreads;
    int lo = s * nthreads + w * 64, hi = s * nthreads + std::min(nthreads, w * 64 + 64);
    bool any = false;
    int op = -1;
    for (int i = lo; i < hi; i++) {
        Fiber& f = x->fibers[i];
        if (f.st == READY) return false;
        if (f.st != AT_WAVE) continue;
        if (!any) {
            op = f.op;
            any = true;
        } else if (f.op != op) {
            for (int q = lo; q < hi; q++)
                fprintf(stderr, "emu:   lane %d state %d op %d\n", q - lo, (int)x->fibers[q].st, x->fibers[q].op);
            die("lanes of one wave wait at different wave operations");
        }
    }
    if (!any) return false;
    auto part = [&](int lane) { return lo + lane < hi && x->fibers[lo + lane].st == AT_WAVE; };
    unsigned long long ballot = 0;
    if (op == 3)
        for (int l = 0; l < hi - lo; l++)
            if (part(l) && x->fibers[lo + l].val) ballot |= 1ull << l;
    for (int l = 0; l < hi - lo; l++) {
        if (!part(l)) continue;
        Fiber& f = x->fibers[lo + l];
        unsigned long long r = f.val;
        if (op == 0) {
            int q = f.arg & 63;
            if (part(q)) r = x->fibers[lo + q].val;
        } else if (op == 1) {
            int q = l - f.arg;
            if (q >= 0 && part(q)) r = x->fibers[lo + q].val;
        } else if (op == 2) {
            int q = l + f.arg;
            if (q < 64 && part(q)) r = x->fibers[lo + q].val;
        } else {
            r = ballot;
        }
        f.res = r;
    }
    for (int l = 0; l < hi - lo; l++)
        if (part(l)) x->fibers[lo + l].st = READY;
    return true;
}


static void start_block(Exec* x, int s, unsigned b)
{
    x->slots[s].bid = U3{b, 0, 0};
    x->slots[s].active = true;
    for (int i = 0; i < x->nthreads; i++) {
        Fiber& f = x->fibers[s * x->nthreads + i];
        fiber_prepare(f);
        f.tid = U3{(unsigned)i, 0, 0};
        f.slot = s;
    }
}

static thread_local Exec t_exec_store;
static unsigned long long g_launches = 0;

void run_grid(const char* name, const LaunchCfg& cfg, Body& body)
{
    std::lock_guard<std::recursive_mutex> g(g_launch_mu);
    if (cfg.grid.y != 1 || cfg.grid.z != 1 || cfg.block.y != 1 || cfg.block.z != 1)
        die("only one-dimensional launches are emulated");
    if (cfg.grid.x == 0 || cfg.block.x == 0 || cfg.block.x > 1024) {
        fprintf(stderr, "emu: invalid launch of %s: grid %u block %u\n", name, cfg.grid.x, cfg.block.x);
        fflush(stderr);
        _exit(90);
    }
    g_launches++;
    Exec* prev = t_exec;
    Exec* x = &t_exec_store;
    if (prev == x) die("a launch from inside a kernel");
    const unsigned G = cfg.grid.x;
    const int nt = (int)cfg.block.x;
    const int C = (int)std::min<unsigned>(G, (unsigned)std::max(1, g_conc));
    x->name = name;
    x->body = &body;
    x->gdim = U3{G, 1, 1};
    x->bdim = U3{(unsigned)nt, 1, 1};
    x->nthreads = nt;
    if ((int)x->fibers.size() < C * nt) x->fibers.resize((size_t)C * nt);
    x->slots.assign(C, Slot{});
    t_exec = x;
    std::vector<unsigned> blocks(G);
    for (unsigned i = 0; i < G; i++) blocks[i] = i;
    for (unsigned i = G - 1; i > 0 && G > 1; i--) std::swap(blocks[i], blocks[splitmix(t_order_state) % (i + 1)]);
    size_t next = 0;
    for (int s = 0; s < C; s++) start_block(x, s, blocks[next++]);
    const int N = C * nt, nw = (nt + 63) / 64;
    std::vector<int> order(N);
    for (int i = 0; i < N; i++) order[i] = i;
    for (;;) {
        // a shuffled order of the work-items between switch points
        for (int i = N - 1; i > 0; i--) std::swap(order[i], order[splitmix(t_order_state) % (i + 1)]);
        for (int k = 0; k < N; k++) {
            Fiber& f = x->fibers[order[k]];
            if (!x->slots[f.slot].active || f.st != READY) continue;
            if (f.stall > 0) {
                f.stall--;
                continue;
            }
            x->cur = &f;
            emu_ctx_switch(&x->sched_sp, f.sp);
            x->cur = nullptr;
        }
        bool any_active = false, progress = false;
        for (int s = 0; s < C; s++) {
            if (!x->slots[s].active) continue;
            for (int w = 0; w < nw; w++) progress |= release_wave(x, s, w);
            bool live = false, all_bar = true, ready = false;
            for (int i = 0; i < nt; i++) {
                State st = x->fibers[s * nt + i].st;
                if (st == DONE) continue;
                live = true;
                if (st == READY) ready = true;
                if (st != AT_BARRIER) all_bar = false;
            }
            if (ready) progress = true;
            if (!live) {
                if (next < G) start_block(x, s, blocks[next++]);
                else x->slots[s].active = false;
                progress = true;
            } else if (all_bar) {
                for (int i = 0; i < nt; i++)
                    if (x->fibers[s * nt + i].st == AT_BARRIER) x->fibers[s * nt + i].st = READY;
                progress = true;
            }
            any_active |= x->slots[s].active;
        }
        if (!any_active) break;
        if (!progress) die("work-items blocked without a barrier or wave operation to complete");
    }
    t_exec = prev;
}

// ------------------------------------------------------------------ faults

static void on_fault(int sig, siginfo_t* si, void*)
{
    char buf[512];
    Exec* x = t_exec;
    uintptr_t a = (uintptr_t)si->si_addr;
    int n = snprintf(buf, sizeof buf, "emu: %s at %p in kernel %s, block %u, thread %u\n",
                     sig == SIGSEGV ? "SIGSEGV" : "SIGBUS", si->si_addr, x ? x->name : "(host code)",
                     x && x->cur ? x->slots[x->cur->slot].bid.x : 0u, x && x->cur ? x->cur->tid.x : 0u);
    (void)!write(2, buf, n);
    // the allocation whose mapping holds the address (guard pages included)
    for (auto& kv : g_allocs) {
        const Alloc& al = kv.second;
        if (a >= (uintptr_t)al.base && a < (uintptr_t)al.base + al.map_bytes) {
            long long off = (long long)(a - (uintptr_t)al.user);
            n = snprintf(buf, sizeof buf, "emu:   allocation #%llu of %zu bytes at %p: offset %lld\n", al.id, al.bytes,
                         (void*)al.user, off);
            (void)!write(2, buf, n);
        }
    }
    _exit(86);
}

struct Init {
    Init()
    {
        if (const char* e = getenv("EMU_SEED")) g_seed = strtoull(e, nullptr, 10);
        if (const char* e = getenv("EMU_ALIGN_LOW")) g_align_low = atoi(e);
        if (const char* e = getenv("EMU_CONC")) g_conc = std::max(1, atoi(e));
        if (const char* e = getenv("EMU_YIELD_ATOMICS")) g_yield_atomics = atoi(e);
        if (const char* e = getenv("EMU_YIELD_MEM")) g_yield_mem = atoi(e);
        t_order_state = g_seed * 0x9E3779B97F4A7C15ull + 1;
        g_fill_state ^= g_seed;
        struct sigaction sa;
        memset(&sa, 0, sizeof sa);
        sa.sa_sigaction = on_fault;
        sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
        static char altstack[65536];
        stack_t ss;
        ss.ss_sp = altstack;
        ss.ss_size = sizeof altstack;
        ss.ss_flags = 0;
        sigaltstack(&ss, nullptr);
        sigaction(SIGSEGV, &sa, nullptr);
        sigaction(SIGBUS, &sa, nullptr);
    }
};
static Init g_init;

}  // namespace emu

// ------------------------------------------------------------------ runtime API

struct emu_stream_t {
    int dummy;
};
struct emu_event_t {
    std::chrono::steady_clock::time_point t;
};

hipError_t hipMalloc(void** p, size_t bytes)
{
    *p = emu::mem_alloc(bytes, true);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p)
{
    emu::mem_free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned)
{
    *p = emu::mem_alloc(bytes, true);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p)
{
    emu::mem_free(p);
    return hipSuccess;
}
hipError_t hipHostGetDevicePointer(void** d, void* h, unsigned)
{
    *d = h;
    return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind)
{
    if (bytes) memmove(dst, src, bytes);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind k, hipStream_t)
{
    return hipMemcpy(dst, src, bytes, k);
}
hipError_t hipMemset(void* dst, int v, size_t bytes)
{
    if (bytes) memset(dst, v, bytes);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t) { return hipMemset(dst, v, bytes); }
hipError_t hipStreamCreate(hipStream_t* s)
{
    *s = new emu_stream_t{0};
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s)
{
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamQuery(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = new emu_event_t{std::chrono::steady_clock::now()};
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t)
{
    e->t = std::chrono::steady_clock::now();
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b)
{
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipPeekAtLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated error"; }
hipError_t hipGetDevice(int* d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceCount(int* n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int)
{
    memset(p, 0, sizeof *p);
    strcpy(p->name, "cpu-simt-emulator");
    strcpy(p->gcnArchName, "emu");
    p->multiProcessorCount = 1;
    p->totalGlobalMem = size_t(64) << 30;
    return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* f, size_t* t)
{
    *f = *t = size_t(64) << 30;
    return hipSuccess;
}

// ------------------------------------------------------------------ pool

// Exact-size allocations instead of size classes, so an access past the
// requested size faults; a release must name a size of the request's class.
static std::mutex g_pool_mu;
static size_t g_in_use = 0, g_peak = 0;

void* MemCache::getMemByAbtrtSize(size_t size)
{
    void* p = emu::mem_alloc(size, true);
    std::lock_guard<std::mutex> g(g_pool_mu);
    g_in_use += classBytes(size);
    g_peak = std::max(g_peak, g_in_use);
    inUse_ = g_in_use;
    peak_ = g_peak;
    return p;
}

void MemCache::releaseAbtrtMem(void* ptr, size_t size)
{
    if (!ptr) return;
    size_t req = emu::mem_size(ptr);
    if (req == 0) {
        fprintf(stderr, "emu: pool release of %p, which is not an allocation\n", ptr);
        fflush(stderr);
        _exit(87);
    }
    if (classBytes(size) != classBytes(req)) {
        fprintf(stderr, "emu: pool release of %p with %zu bytes; it was requested with %zu (another size class)\n", ptr,
                size, req);
        fflush(stderr);
        _exit(91);
    }
    emu::mem_free(ptr);
    std::lock_guard<std::mutex> g(g_pool_mu);
    g_in_use -= std::min(g_in_use, classBytes(size));
    inUse_ = g_in_use;
}

static const size_t kSizes[] = {32, 1024, 65536, 131072, 262144, 524288, 1048576, 2097152,
                                4194304, 8388608, 25165824, 536870912};
size_t MemCache::classBytes(size_t size)
{
    for (size_t s : kSizes)
        if (size <= s) return s;
    const size_t step = size_t(64) << 20;
    return (size + step - 1) / step * step;
}
void* MemCache::getMemBySize(size_t) { return nullptr; }
void MemCache::releaseMem(void*, size_t) {}
void MemCache::initAbtrtMem() {}
void MemCache::releaseCache() {}
void* MemCache::getPinnedPage() { return m_pinnedPage; }
void MemCache::releasePinnedPage(void* mem) { m_pinnedPage = mem; }
void MemCache::report() {}
MemCache& getMemCacheRef()
{
    static MemCache cache;
    return cache;
}

// ------------------------------------------------------------------ test hooks

extern "C" {
// Sets a round tunable without the GPU-coverage rule (grid_cap may go below
// 256 here, so grid-stride loops run many iterations on small inputs).
int emu_set_tunable(const char* name, int value)
{
    std::string n(name);
    auto& t = gcdt::Tunables::get();
    if (n == "tail_threshold") t.tail_threshold = value;
    else if (n == "pivot") t.pivot = value;
    else if (n == "single_tester") t.single_tester = value;
    else if (n == "grid_cap") gcdt::g_grid_cap = value;
    else if (n == "flip_items") t.flip_items = value;
    else if (n == "ins_items") t.ins_items = value;
    else if (n == "validate") t.validate = value;
    else return 2;
    return 0;
}
unsigned long long emu_launches(void) { return emu::g_launches; }
}

// Self-tests of the detectors: a kernel that writes one element past the end
// (where > 0) or before the start (where < 0) of a 1000-int allocation, and a
// pool release that names another size class.
static void emu_selftest(int* p, int n, int where)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) p[where > 0 ? n : -1] = 1;
}
extern "C" void emu_selftest_oob(int where)
{
    int* p = nullptr;
    (void)hipMalloc(&p, sizeof(int) * 1000);
    emu::launch("emu_selftest", emu::LaunchCfg(dim3(2), dim3(64)), &emu_selftest, p, 1000, where);
}
extern "C" void emu_selftest_release(void)
{
    void* p = getMemCacheRef().getMemByAbtrtSize(5000);
    getMemCacheRef().releaseAbtrtMem(p, 100000);
}
