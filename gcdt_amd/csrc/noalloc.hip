// gcdt_solve without device allocation: every buffer of the solve is carved
// out of one static device array (`scratch`) of fixed size, for callers that
// cannot call hipMalloc (graph capture, serving loops, harnesses that forbid
// it).  The caller passes inputs and outputs as for gcdt_solve.
//
// Set-up happens in gcdt_pool_init(), once per device with that device
// current: the array's device address is looked up with
// hipGetSymbolAddress (the array's name is a host-side stand-in that no
// kernel may be given) and a page of pinned host counters is reserved for
// this path.  The solve itself then allocates nothing: its buffers come from
// the array through a bump arena (mesh.hpp, Arena), the counters from the
// reserved page, and its timings from the host clock instead of events.
//
// Every call on a device shares that device's array and keeps the last call's
// contents; the solve initialises everything it reads.  Calls on one device
// are serialised by a lock (each call returns after its stream has drained),
// so overlapping calls wait instead of racing on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mesh.hpp"

using namespace gcdt;

namespace gcdt {
int fail(int status, const std::string& msg);
}

namespace {

// The solve's live buffers peak near 40 MB + 260 bytes per point (measured),
// plus about 48 bytes per triangle crossed by a constraint segment.
constexpr size_t kScratchBytes = GCDT_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceState {
    void* ptr = nullptr;  // the device address of `scratch` on this device
    HostSlot slot;        // pinned counter pages reserved for this path
    bool ready = false;
    std::mutex call;      // one call at a time per device
};
DeviceState g_dev[kMaxDevices];
std::mutex g_setup_mu;

}  // namespace

namespace gcdt {

// Called by gcdt_pool_init with the device current.
void noalloc_setup()
{
    int dev = 0;
    GCDT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(GCDT_ERR_INVALID, "device id out of range");
    std::lock_guard<std::mutex> g(g_setup_mu);
    DeviceState& d = g_dev[dev];
    if (d.ready) return;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(GCDT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    d.ptr = p;
    d.slot = take_host_slot();
    d.ready = true;
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

}  // namespace gcdt

extern "C" int gcdt_solve_noalloc(const gcdt_int2* points, int N, const gcdt_int2* cons, int C, gcdt_int3* tris_out,
                                  gcdt_int3* adj_out, int* num_tris, gcdt_stats* stats, void* stream)
{
    try {
        if (N < 0 || C < 0 || (N > 0 && !points) || (C > 0 && !cons) || !tris_out)
            return fail(GCDT_ERR_INVALID, "bad arguments");
        if (N > GCDT_NOALLOC_MAX_POINTS || C > GCDT_NOALLOC_MAX_POINTS)
            return fail(GCDT_ERR_SCRATCH, "gcdt_solve_noalloc takes at most " +
                                              std::to_string(GCDT_NOALLOC_MAX_POINTS) +
                                              " points and as many segments (its scratch array is fixed); use "
                                              "gcdt_solve");
        int dev = 0;
        GCDT_HIP(hipGetDevice(&dev));
        if (dev < 0 || dev >= kMaxDevices || !g_dev[dev].ready)
            return fail(GCDT_ERR_INVALID, "gcdt_pool_init() has not run on this device: it sets up the scratch "
                                          "array of gcdt_solve_noalloc");
        DeviceState& d = g_dev[dev];
        std::lock_guard<std::mutex> lock(d.call);
        // the stream drains before the lock is released, also after an error,
        // so no kernel of this call still uses the array when the next starts
        struct Drain {
            hipStream_t s;
            ~Drain() { (void)hipStreamSynchronize(s); }
        } drain{(hipStream_t)stream};
        Arena arena(d.ptr, kScratchBytes);
        struct Use {
            Use(Arena* a) { g_arena = a; }
            ~Use() { g_arena = nullptr; }
        } use(&arena);
        SolveOut out;
        {
            Ctx c((hipStream_t)stream, &d.slot);
            DBuf<int3> adj_tmp;
            int3* adj = (int3*)adj_out;
            if (!adj) {
                adj_tmp.alloc(2LL * N + 2);
                adj = adj_tmp.p;
            }
            solve_impl(c, (const int2*)points, N, (const int2*)cons, C, (int3*)tris_out, adj, out);
        }
        if (num_tris) *num_tris = out.T;
        if (stats) *stats = out.st;
        return GCDT_OK;
    } catch (const Error& e) {
        return fail(e.status, e.what());
    } catch (const std::exception& e) {
        return fail(GCDT_ERR_INTERNAL, e.what());
    }
}
