// Even-odd carving of a constrained triangulation: the triangles inside the
// constraint polygons (holes removed), where the reference triangulates the
// whole bounding square and keeps everything.
//
// 1. Connected components of the triangles across unconstrained edges:
//    min-label hooking (atomicMin on roots) and pointer jumping, repeated
//    until no edge joins two components.
// 2. The component graph (one edge per constrained mesh edge between two
//    components) is small -- about one node per polygon -- and is walked
//    breadth-first on the host from the component that holds the square's
//    corners; a component's depth is the fewest constraint edges crossed to
//    reach it from outside, and odd depth means inside.
#include <algorithm>
#include <deque>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/gcdt.h"
#include "mesh.hpp"

namespace gcdt {
int fail(int status, const std::string& msg);
}

using namespace gcdt;

namespace {

__device__ inline int cc_find(const int* label, int x)
{
    int p = label[x];
    while (p != x) {
        x = p;
        p = label[x];
    }
    return x;
}

__global__ __launch_bounds__(kBlock) void k_cc_init(int* label, int T)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x)
        label[i] = (int)i;
}

// Join the roots of both sides of every unconstrained edge (each edge once,
// from its lower triangle): the larger root takes the smaller as parent.
__global__ __launch_bounds__(kBlock) void k_cc_hook(const int3* adj, const int* cflag, int T, int* label, int* changed)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        int t = (int)i;
        int3 a = adj[t];
        int cf = cflag ? cflag[t] : 0;
        for (int k = 0; k < 3; k++) {
            int n = get3(a, k);
            if (n <= t || ((cf >> k) & 1)) continue;
            int r1 = cc_find(label, t), r2 = cc_find(label, n);
            if (r1 == r2) continue;
            atomicMin(&label[max(r1, r2)], min(r1, r2));
            *changed = 1;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_cc_compress(int* label, int T)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x)
        label[i] = cc_find(label, (int)i);
}

// Constrained mesh edges between two different components, each edge once.
__global__ __launch_bounds__(kBlock) void k_carve_pairs(const int3* adj, const int* cflag, const int* label, int T,
                                                        int2* pairs, int cap, int* npairs)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        int t = (int)i;
        int cf = cflag[t];
        if (!cf) continue;
        int3 a = adj[t];
        for (int k = 0; k < 3; k++) {
            int n = get3(a, k);
            if (n <= t || !((cf >> k) & 1)) continue;
            int ra = label[t], rb = label[n];
            if (ra == rb) continue;
            int o = atomicAdd(npairs, 1);
            if (o < cap) pairs[o] = make_int2(ra, rb);
        }
    }
}

// Lowest triangle that has a corner of the square (vertex id >= n).
__global__ __launch_bounds__(kBlock) void k_corner_tri(const int3* tri, int T, int n, int* out)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        int3 v = tri[i];
        if (v.x >= n || v.y >= n || v.z >= n) atomicMin(out, (int)i);
    }
}

__global__ __launch_bounds__(kBlock) void k_root_flags(const int* roots, int nr, int* rootflag)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += (long long)gridDim.x * blockDim.x)
        rootflag[roots[i]] = 1;
}

__global__ __launch_bounds__(kBlock) void k_inside(const int* label, const int* rootflag, int T, int* inside, int* count)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x) {
        int f = rootflag[label[i]];
        inside[i] = f;
        if (f) atomicAdd(count, 1);
    }
}

}  // namespace

extern "C" int gcdt_carve(const gcdt_int3* tris, const gcdt_int3* adj, const int* cflag, int T, int numPoints,
                          int* inside, int* num_inside, void* stream)
{
    try {
        if (T <= 0 || !tris || !adj || !cflag || !inside || numPoints < 0)
            return fail(GCDT_ERR_INVALID, "bad arguments");
        Ctx c((hipStream_t)stream);
        DBuf<int> label(T);
        k_cc_init<<<grid_for(T), kBlock, 0, c.st>>>(label.p, T);
        GCDT_LAUNCH_CHECK();
        for (int it = 0;; it++) {
            if (it > 4096) throw Error(GCDT_ERR_NOT_CONVERGED, "component labelling did not converge");
            c.set(C_TOTAL, 0);
            k_cc_hook<<<grid_for(T), kBlock, 0, c.st>>>((const int3*)adj, cflag, T, label.p, c.dcnt.p + C_TOTAL);
            GCDT_LAUNCH_CHECK();
            k_cc_compress<<<grid_for(T), kBlock, 0, c.st>>>(label.p, T);
            GCDT_LAUNCH_CHECK();
            if (!c.get(C_TOTAL)) break;
        }
        // component graph
        c.set(C_TOTAL, 0);
        k_carve_pairs<<<grid_for(T), kBlock, 0, c.st>>>((const int3*)adj, cflag, label.p, T, nullptr, 0,
                                                         c.dcnt.p + C_TOTAL);
        GCDT_LAUNCH_CHECK();
        int np = c.get(C_TOTAL);
        std::vector<int2> pairs(np);
        if (np > 0) {
            DBuf<int2> dpairs(np);
            c.set(C_TOTAL, 0);
            k_carve_pairs<<<grid_for(T), kBlock, 0, c.st>>>((const int3*)adj, cflag, label.p, T, dpairs.p, np,
                                                             c.dcnt.p + C_TOTAL);
            GCDT_LAUNCH_CHECK();
            GCDT_HIP(hipMemcpyAsync(pairs.data(), dpairs.p, sizeof(int2) * np, hipMemcpyDeviceToHost, c.st));
            GCDT_HIP(hipStreamSynchronize(c.st));
        }
        c.set(C_TOTAL, T);
        k_corner_tri<<<grid_for(T), kBlock, 0, c.st>>>((const int3*)tris, T, numPoints, c.dcnt.p + C_TOTAL);
        GCDT_LAUNCH_CHECK();
        int t0 = c.get(C_TOTAL);
        if (t0 >= T) return fail(GCDT_ERR_INVALID, "no triangle has a corner of the square");
        int outer = 0;
        GCDT_HIP(hipMemcpyAsync(&outer, label.p + t0, sizeof(int), hipMemcpyDeviceToHost, c.st));
        GCDT_HIP(hipStreamSynchronize(c.st));
        // breadth-first depth over the component graph
        std::unordered_map<int, std::vector<int>> g;
        for (const int2& e : pairs) {
            g[e.x].push_back(e.y);
            g[e.y].push_back(e.x);
        }
        std::unordered_map<int, int> depth;
        depth[outer] = 0;
        std::deque<int> q{outer};
        std::vector<int> odd;
        while (!q.empty()) {
            int u = q.front();
            q.pop_front();
            auto it = g.find(u);
            if (it == g.end()) continue;
            for (int w : it->second) {
                if (depth.count(w)) continue;
                depth[w] = depth[u] + 1;
                if (depth[w] & 1) odd.push_back(w);
                q.push_back(w);
            }
        }
        DBuf<int> rootflag(T);
        fill_i32(c, rootflag.p, 0, T);
        if (!odd.empty()) {
            DBuf<int> droots(odd.size());
            GCDT_HIP(hipMemcpyAsync(droots.p, odd.data(), sizeof(int) * odd.size(), hipMemcpyHostToDevice, c.st));
            k_root_flags<<<grid_for((long long)odd.size()), kBlock, 0, c.st>>>(droots.p, (int)odd.size(), rootflag.p);
            GCDT_LAUNCH_CHECK();
            GCDT_HIP(hipStreamSynchronize(c.st));  // the host vector is the copy's source
        }
        c.set(C_TOTAL, 0);
        k_inside<<<grid_for(T), kBlock, 0, c.st>>>(label.p, rootflag.p, T, inside, c.dcnt.p + C_TOTAL);
        GCDT_LAUNCH_CHECK();
        int cnt = c.get(C_TOTAL);
        if (num_inside) *num_inside = cnt;
        return GCDT_OK;
    } catch (const std::bad_alloc&) {
        return fail(GCDT_ERR_OOM, "host allocation failed");
    } catch (const Error& e) {
        return fail(e.status, e.what());
    }
}
