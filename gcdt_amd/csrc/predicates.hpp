// Exact integer predicates on the 2^30 point grid.
//
// orient2d follows include/math_base.cuh:68-76 of the reference: the sign of
// (a-c) x (b-c), +1 counter-clockwise, 0 collinear, -1 clockwise.  Coordinate
// differences are < 2^30 in magnitude, so both products fit in int64.
//
// incircle follows include/math_base.cuh:267-328: the sign of the lifted
// determinant | a-p |a-p|^2 ; b-p |b-p|^2 ; c-p |c-p|^2 |, +1 when p lies
// strictly inside the circumcircle of counter-clockwise a,b,c.  It is
// evaluated by cofactor expansion along the lifted column:
//     det = |a-p|^2 * X(b,c) + |b-p|^2 * X(c,a) + |c-p|^2 * X(a,b)
// with X(u,v) = u.x*v.y - u.y*v.x (|X| < 2^61 exactly in int64) and
// |u-p|^2 < 2^61.  Each product needs < 2^122 bits, the sum < 2^124, so the
// whole determinant is exact in a signed 128-bit value built here from two
// 64-bit limbs (no floating point at any step).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gcdt {

struct I128 {
    uint64_t lo;
    int64_t hi;
};

// Full signed 64x64 -> 128 product.
__device__ __host__ inline I128 mul_s64(int64_t a, int64_t b)
{
    I128 r;
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t ua = (uint64_t)a, ub = (uint64_t)b;
    uint64_t hi = __umul64hi(ua, ub);
    // signed correction of the unsigned high word
    if (a < 0) hi -= ub;
    if (b < 0) hi -= ua;
    r.lo = ua * ub;
    r.hi = (int64_t)hi;
#else
    __int128 p = (__int128)a * (__int128)b;
    r.lo = (uint64_t)p;
    r.hi = (int64_t)(p >> 64);
#endif
    return r;
}

__device__ __host__ inline I128 add128(I128 a, I128 b)
{
    I128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1 : 0);
    return r;
}

__device__ __host__ inline int sign128(I128 a)
{
    if (a.hi < 0) return -1;
    if (a.hi > 0) return 1;
    return a.lo != 0 ? 1 : 0;
}

__device__ __host__ inline int orient2d(int2 a, int2 b, int2 c)
{
    int64_t o1 = (int64_t)(a.x - c.x) * (int64_t)(b.y - c.y);
    int64_t o2 = (int64_t)(a.y - c.y) * (int64_t)(b.x - c.x);
    return o1 > o2 ? 1 : (o1 < o2 ? -1 : 0);
}

__device__ __host__ inline int incircle(int2 a, int2 b, int2 c, int2 p)
{
    int64_t ax = a.x - p.x, ay = a.y - p.y;
    int64_t bx = b.x - p.x, by = b.y - p.y;
    int64_t cx = c.x - p.x, cy = c.y - p.y;
    int64_t aa = ax * ax + ay * ay;
    int64_t bb = bx * bx + by * by;
    int64_t cc = cx * cx + cy * cy;
    int64_t xbc = bx * cy - by * cx;
    int64_t xca = cx * ay - cy * ax;
    int64_t xab = ax * by - ay * bx;
    I128 s = add128(add128(mul_s64(aa, xbc), mul_s64(bb, xca)), mul_s64(cc, xab));
    return sign128(s);
}

// p strictly between a and b on the line through them (caller knows collinear)
__device__ __host__ inline bool strictly_between(int2 a, int2 b, int2 p)
{
    int64_t d1 = (int64_t)(p.x - a.x) * (b.x - a.x) + (int64_t)(p.y - a.y) * (b.y - a.y);
    int64_t d2 = (int64_t)(p.x - b.x) * (a.x - b.x) + (int64_t)(p.y - b.y) * (a.y - b.y);
    return d1 > 0 && d2 > 0;
}

}  // namespace gcdt
