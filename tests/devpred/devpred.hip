// Test-only kernels: the library's exact predicates (gcdt_amd/csrc/
// predicates.hpp) evaluated on the GPU over caller-given point tuples, so the
// device code path -- the 64x64->128 product from __umul64hi with its signed
// correction, which host builds replace by __int128 -- is checked against the
// golden vectors directly.  Not part of the library.
#include <hip/hip_runtime.h>

#include "../../gcdt_amd/csrc/exact.hpp"
#include "../../gcdt_amd/csrc/predicates.hpp"

namespace {

__global__ void k_orient(const int2* p, int n, int* out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = gcdt::orient2d(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

__global__ void k_incircle(const int2* p, int n, int* out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = gcdt::incircle(p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]);
}

// the signed 128-bit product's two limbs, for products checked against
// Python integers
__global__ void k_mul(const long long* ab, int n, unsigned long long* lo, long long* hi)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        gcdt::I128 r = gcdt::mul_s64(ab[2 * i], ab[2 * i + 1]);
        lo[i] = r.lo;
        hi[i] = r.hi;
    }
}

// homogeneous points as 5 int64 words: X low, X high, Y low, Y high, W
__device__ gcdt::HPt hpt_at(const long long* q)
{
    gcdt::HPt h;
    h.X.w[0] = (uint64_t)q[0];
    h.X.w[1] = (uint64_t)q[1];
    h.Y.w[0] = (uint64_t)q[2];
    h.Y.w[1] = (uint64_t)q[3];
    h.W = q[4];
    return h;
}

__global__ void k_orient_h(const long long* p, int n, int* out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = gcdt::orient_h(hpt_at(p + 15 * i), hpt_at(p + 15 * i + 5), hpt_at(p + 15 * i + 10));
}

__global__ void k_incircle_h(const long long* p, int n, int* out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = gcdt::incircle_h(hpt_at(p + 20 * i), hpt_at(p + 20 * i + 5), hpt_at(p + 20 * i + 10),
                                  hpt_at(p + 20 * i + 15));
}

// crossing of two grid segments: 8 ints in, 5 words out (W 0: parallel)
__global__ void k_cross(const int* s, int n, long long* out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int* q = s + 8 * i;
        gcdt::HPt h;
        long long* o = out + 5 * i;
        if (gcdt::hpt_cross(make_int2(q[0], q[1]), make_int2(q[2], q[3]), make_int2(q[4], q[5]),
                            make_int2(q[6], q[7]), h)) {
            o[0] = (long long)h.X.w[0];
            o[1] = (long long)h.X.w[1];
            o[2] = (long long)h.Y.w[0];
            o[3] = (long long)h.Y.w[1];
            o[4] = h.W;
        } else {
            for (int k = 0; k < 5; k++) o[k] = 0;
        }
    }
}

int run(const void* in, size_t in_bytes, int n, int kind, void* out0, size_t out0_bytes, void* out1)
{
    void *din = nullptr, *d0 = nullptr, *d1 = nullptr;
    int rc = 0;
    if (hipMalloc(&din, in_bytes ? in_bytes : 1) != hipSuccess || hipMalloc(&d0, out0_bytes ? out0_bytes : 1) != hipSuccess ||
        hipMalloc(&d1, out0_bytes ? out0_bytes : 1) != hipSuccess) {
        rc = 1;
    } else if (n > 0) {
        hipMemcpy(din, in, in_bytes, hipMemcpyHostToDevice);
        int g = (n + 255) / 256;
        if (kind == 0) k_orient<<<g, 256>>>((const int2*)din, n, (int*)d0);
        else if (kind == 1) k_incircle<<<g, 256>>>((const int2*)din, n, (int*)d0);
        else if (kind == 2) k_mul<<<g, 256>>>((const long long*)din, n, (unsigned long long*)d0, (long long*)d1);
        else if (kind == 3) k_orient_h<<<g, 256>>>((const long long*)din, n, (int*)d0);
        else if (kind == 4) k_incircle_h<<<g, 256>>>((const long long*)din, n, (int*)d0);
        else k_cross<<<g, 256>>>((const int*)din, n, (long long*)d0);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = 2;
        else {
            hipMemcpy(out0, d0, out0_bytes, hipMemcpyDeviceToHost);
            if (out1) hipMemcpy(out1, d1, out0_bytes, hipMemcpyDeviceToHost);
        }
    }
    hipFree(din);
    hipFree(d0);
    hipFree(d1);
    return rc;
}

}  // namespace

// xy: n tuples of 3 (orient2d) or 4 (incircle) points as int32 x,y pairs
extern "C" int devpred_orient2d(const int* xy, int n, int* out)
{
    return run(xy, (size_t)n * 6 * sizeof(int), n, 0, out, (size_t)n * sizeof(int), nullptr);
}

extern "C" int devpred_incircle(const int* xy, int n, int* out)
{
    return run(xy, (size_t)n * 8 * sizeof(int), n, 1, out, (size_t)n * sizeof(int), nullptr);
}

// ab: n pairs of int64 factors; lo/hi: the limbs of each signed product
extern "C" int devpred_mul_s64(const long long* ab, int n, unsigned long long* lo, long long* hi)
{
    return run(ab, (size_t)n * 2 * sizeof(long long), n, 2, lo, (size_t)n * 8, hi);
}

// homogeneous predicates (exact.hpp): pts = n tuples of 3 (orient) or 4
// (incircle) points of 5 int64 words each
extern "C" int devpred_orient_h(const long long* pts, int n, int* out)
{
    return run(pts, (size_t)n * 15 * sizeof(long long), n, 3, out, (size_t)n * sizeof(int), nullptr);
}

extern "C" int devpred_incircle_h(const long long* pts, int n, int* out)
{
    return run(pts, (size_t)n * 20 * sizeof(long long), n, 4, out, (size_t)n * sizeof(int), nullptr);
}

// segs: n pairs of grid segments (8 ints); out: 5 words per crossing
extern "C" int devpred_cross(const int* segs, int n, long long* out)
{
    return run(segs, (size_t)n * 8 * sizeof(int), n, 5, out, (size_t)n * 5 * sizeof(long long), nullptr);
}
