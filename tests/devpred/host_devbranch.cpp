// The device branch of the library's predicates (gcdt_amd/csrc/predicates.hpp)
// compiled for the host: the emulator's stand-in <hip/hip_runtime.h>
// (tests/emu/hip) defines __HIP_DEVICE_COMPILE__ and an exact __umul64hi, so
// mul_s64 takes its GPU path -- the unsigned high word with the signed
// correction -- instead of __int128.  A CPU test of that path (test
// infrastructure, not product code).
#include <hip/hip_runtime.h>

#include "../../gcdt_amd/csrc/exact.hpp"
#include "../../gcdt_amd/csrc/predicates.hpp"

extern "C" int hostdev_orient2d(const int* xy, int n, int* out)
{
    for (int i = 0; i < n; i++) {
        const int* q = xy + 6 * i;
        out[i] = gcdt::orient2d(make_int2(q[0], q[1]), make_int2(q[2], q[3]), make_int2(q[4], q[5]));
    }
    return 0;
}

extern "C" int hostdev_incircle(const int* xy, int n, int* out)
{
    for (int i = 0; i < n; i++) {
        const int* q = xy + 8 * i;
        out[i] = gcdt::incircle(make_int2(q[0], q[1]), make_int2(q[2], q[3]), make_int2(q[4], q[5]),
                                make_int2(q[6], q[7]));
    }
    return 0;
}

extern "C" int hostdev_mul_s64(const long long* ab, int n, unsigned long long* lo, long long* hi)
{
    for (int i = 0; i < n; i++) {
        gcdt::I128 r = gcdt::mul_s64(ab[2 * i], ab[2 * i + 1]);
        lo[i] = r.lo;
        hi[i] = r.hi;
    }
    return 0;
}

static gcdt::HPt hpt_at(const long long* q)
{
    gcdt::HPt h;
    h.X.w[0] = (uint64_t)q[0];
    h.X.w[1] = (uint64_t)q[1];
    h.Y.w[0] = (uint64_t)q[2];
    h.Y.w[1] = (uint64_t)q[3];
    h.W = q[4];
    return h;
}

extern "C" int hostdev_orient_h(const long long* p, int n, int* out)
{
    for (int i = 0; i < n; i++) out[i] = gcdt::orient_h(hpt_at(p + 15 * i), hpt_at(p + 15 * i + 5), hpt_at(p + 15 * i + 10));
    return 0;
}

extern "C" int hostdev_incircle_h(const long long* p, int n, int* out)
{
    for (int i = 0; i < n; i++)
        out[i] = gcdt::incircle_h(hpt_at(p + 20 * i), hpt_at(p + 20 * i + 5), hpt_at(p + 20 * i + 10),
                                  hpt_at(p + 20 * i + 15));
    return 0;
}

extern "C" int hostdev_cross(const int* s, int n, long long* out)
{
    for (int i = 0; i < n; i++) {
        const int* q = s + 8 * i;
        gcdt::HPt h;
        long long* o = out + 5 * i;
        if (gcdt::hpt_cross(make_int2(q[0], q[1]), make_int2(q[2], q[3]), make_int2(q[4], q[5]), make_int2(q[6], q[7]),
                            h)) {
            o[0] = (long long)h.X.w[0];
            o[1] = (long long)h.X.w[1];
            o[2] = (long long)h.Y.w[0];
            o[3] = (long long)h.Y.w[1];
            o[4] = h.W;
        } else {
            for (int k = 0; k < 5; k++) o[k] = 0;
        }
    }
    return 0;
}
