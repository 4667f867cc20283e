// Exact predicates on points with rational coordinates (host and device).
//
// A crossing of two constraint segments whose endpoints lie on the 2^30 grid
// is a rational point: with d1 = P2 - P1, d2 = Q2 - Q1 and
// W = d1 x d2 (|W| < 2^61), t = (Q1 - P1) x d2, the crossing is
// (X/W, Y/W) with X = P1.x W + d1.x t, Y = P1.y W + d1.y t (|X|, |Y| < 2^92).
// Such points are kept in homogeneous form (X, Y, W), W > 0, and a grid
// point is (x, y, 1).  The predicates are the signs of determinants of the
// homogeneous coordinates, evaluated in fixed-width two's-complement integers
// wide enough for every term:
//   orient(p, q, r) = sign | Xp Yp Wp ; Xq Yq Wq ; Xr Yr Wr |         (< 2^248: 4 limbs)
//   incircle(a, b, c, d) = sign | XW  YW  X^2+Y^2  W^2 |_{a,b,c,d}   (< 2^618: 10 limbs)
// which equal orient2d / incircle of predicates.hpp on grid points (rows of
// the second are scaled by W^2 > 0, which keeps the sign).  No floating point.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "predicates.hpp"

namespace gcdt {

// 64 x 64 -> 128 unsigned product: low word returned, high word in *hi.
__host__ __device__ inline uint64_t mulu64(uint64_t a, uint64_t b, uint64_t* hi)
{
#if defined(__HIP_DEVICE_COMPILE__)
    *hi = __umul64hi(a, b);
    return a * b;
#else
    unsigned __int128 p = (unsigned __int128)a * b;
    *hi = (uint64_t)(p >> 64);
    return (uint64_t)p;
#endif
}

// Signed integer of L 64-bit limbs (little-endian, two's complement).
template <int L>
struct Big {
    uint64_t w[L];

    __host__ __device__ static Big of(int64_t v)
    {
        Big r;
        r.w[0] = (uint64_t)v;
        for (int i = 1; i < L; i++) r.w[i] = v < 0 ? ~0ull : 0ull;
        return r;
    }
    // from a narrower value (sign-extended)
    template <int M>
    __host__ __device__ static Big of(const Big<M>& v)
    {
        Big r;
        uint64_t ext = (v.w[M - 1] >> 63) ? ~0ull : 0ull;
        for (int i = 0; i < L; i++) r.w[i] = i < M ? v.w[i] : ext;
        return r;
    }
    __host__ __device__ bool neg() const { return (w[L - 1] >> 63) != 0; }
    __host__ __device__ int sign() const
    {
        if (neg()) return -1;
        for (int i = 0; i < L; i++)
            if (w[i]) return 1;
        return 0;
    }
};

template <int L>
__host__ __device__ inline Big<L> operator+(const Big<L>& a, const Big<L>& b)
{
    Big<L> r;
    uint64_t carry = 0;
    for (int i = 0; i < L; i++) {
        uint64_t s = a.w[i] + carry;
        uint64_t c1 = s < carry;
        r.w[i] = s + b.w[i];
        carry = c1 + (r.w[i] < s);
    }
    return r;
}

template <int L>
__host__ __device__ inline Big<L> operator-(const Big<L>& a)
{
    Big<L> r;
    uint64_t carry = 1;
    for (int i = 0; i < L; i++) {
        r.w[i] = ~a.w[i] + carry;
        carry = carry && r.w[i] == 0;
    }
    return r;
}

template <int L>
__host__ __device__ inline Big<L> operator-(const Big<L>& a, const Big<L>& b)
{
    return a + (-b);
}

// Product truncated to L limbs (exact when |a * b| < 2^(64 L - 1)).
template <int L>
__host__ __device__ inline Big<L> operator*(const Big<L>& a, const Big<L>& b)
{
    bool na = a.neg(), nb = b.neg();
    Big<L> x = na ? -a : a, y = nb ? -b : b;
    Big<L> r;
    for (int i = 0; i < L; i++) r.w[i] = 0;
    for (int i = 0; i < L; i++) {
        if (!x.w[i]) continue;
        uint64_t carry = 0;
        for (int j = 0; i + j < L; j++) {
            uint64_t hi;
            uint64_t lo = mulu64(x.w[i], y.w[j], &hi);
            uint64_t s = r.w[i + j] + lo;
            uint64_t c = s < lo;
            s += carry;
            c += s < carry;
            r.w[i + j] = s;
            carry = hi + c;
        }
    }
    return (na != nb) ? -r : r;
}

// A point in homogeneous integer coordinates (X/W, Y/W), W > 0.
struct HPt {
    Big<2> X, Y;
    int64_t W;
};

__host__ __device__ inline HPt hpt_grid(int2 p)
{
    HPt h;
    h.X = Big<2>::of((int64_t)p.x);
    h.Y = Big<2>::of((int64_t)p.y);
    h.W = 1;
    return h;
}

// The crossing of the lines P1P2 and Q1Q2 (grid points); false when they are
// parallel.
__host__ __device__ inline bool hpt_cross(int2 P1, int2 P2, int2 Q1, int2 Q2, HPt& out)
{
    int64_t d1x = (int64_t)P2.x - P1.x, d1y = (int64_t)P2.y - P1.y;
    int64_t d2x = (int64_t)Q2.x - Q1.x, d2y = (int64_t)Q2.y - Q1.y;
    int64_t W = d1x * d2y - d1y * d2x;
    if (W == 0) return false;
    int64_t ex = (int64_t)Q1.x - P1.x, ey = (int64_t)Q1.y - P1.y;
    int64_t t = ex * d2y - ey * d2x;
    Big<2> X = Big<2>::of((int64_t)P1.x) * Big<2>::of(W) + Big<2>::of(d1x) * Big<2>::of(t);
    Big<2> Y = Big<2>::of((int64_t)P1.y) * Big<2>::of(W) + Big<2>::of(d1y) * Big<2>::of(t);
    if (W < 0) {
        X = -X;
        Y = -Y;
        W = -W;
    }
    out.X = X;
    out.Y = Y;
    out.W = W;
    return true;
}

// orient2d on homogeneous points (the sign convention of orient2d).
__host__ __device__ inline int orient_h(const HPt& p, const HPt& q, const HPt& r)
{
    typedef Big<4> B;
    B xp = B::of(p.X), yp = B::of(p.Y), wp = B::of(p.W);
    B xq = B::of(q.X), yq = B::of(q.Y), wq = B::of(q.W);
    B xr = B::of(r.X), yr = B::of(r.Y), wr = B::of(r.W);
    B d = xp * (yq * wr - yr * wq) - yp * (xq * wr - xr * wq) + wp * (xq * yr - xr * yq);
    return d.sign();
}

// incircle on homogeneous points (the sign convention of incircle: +1 when d
// lies strictly inside the circle through counter-clockwise a, b, c).  Rows
// are built from the products in the comment at the top; |det| < 2^618.
__host__ __device__ inline int incircle_h(const HPt& a, const HPt& b, const HPt& c, const HPt& d)
{
    typedef Big<10> B;
    const HPt* P[4] = {&a, &b, &c, &d};
    B r[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        B X = B::of(P[i]->X), Y = B::of(P[i]->Y), W = B::of(P[i]->W);
        r[i][0] = X * W;
        r[i][1] = Y * W;
        r[i][2] = X * X + Y * Y;
        r[i][3] = W * W;
    }
    // Laplace expansion by complementary 2x2 minors of columns (0,1) and (2,3)
    auto m = [&](int i, int j) { return r[i][0] * r[j][1] - r[j][0] * r[i][1]; };
    auto n = [&](int i, int j) { return r[i][2] * r[j][3] - r[j][2] * r[i][3]; };
    B det = m(0, 1) * n(2, 3) - m(0, 2) * n(1, 3) + m(0, 3) * n(1, 2) + m(1, 2) * n(0, 3) - m(1, 3) * n(0, 2) +
            m(2, 3) * n(0, 1);
    // | x y x^2+y^2 1 | over (a,b,c,d) equals the lifted 3x3 determinant of
    // predicates.hpp (subtract row d, expand along the last column)
    return det.sign();
}

// sign of x(u) - x(p) and of y(u) - y(p)
__host__ __device__ inline int hpt_cmp_x(const HPt& u, const HPt& p)
{
    typedef Big<4> B;
    return (B::of(u.X) * B::of(p.W) - B::of(p.X) * B::of(u.W)).sign();
}
__host__ __device__ inline int hpt_cmp_y(const HPt& u, const HPt& p)
{
    typedef Big<4> B;
    return (B::of(u.Y) * B::of(p.W) - B::of(p.Y) * B::of(u.W)).sign();
}

}  // namespace gcdt
