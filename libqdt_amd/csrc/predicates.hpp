// predicates.hpp -- orient2d / incircle for gfx950.
//
// Stage 1 is an FP64 error-bound filter with the static bounds and the floor
// of libqdt/robust.hxx:281-289 (same expressions, compiled without FMA
// contraction so the bound derivation holds).  When the filter cannot decide,
// the sign is computed exactly.  The exact stage does not use floating-point
// expansions: every coordinate difference is first split error-free into
// (hi, lo) with two_sum, each component is a dyadic m * 2^e, and with the
// smallest exponent E of all components as unit the differences become
// integers.  When they fit in 62 bits (all coordinates of one predicate
// within a factor ~2^9 of each other, the practical case) the determinant is
// evaluated in 128-bit (orient2d) or 256-bit (incircle) two's-complement
// integers.  Otherwise the expanded monomials are summed into a wide
// fixed-point accumulator.
// Both are exact for all finite doubles (a difference that overflows is kept
// as its two terms, see exact_diff), so the sign equals the
// sign of the true determinant, which is what the reference's exact path
// (robust.hxx:297-396) computes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

// Everything here is __host__ __device__ so that the exact arithmetic can
// also be checked on the CPU.
#define QDT_HD __host__ __device__
#define QDT_HD_NOINLINE __host__ __device__ inline __attribute__((noinline))

namespace qdt
{

QDT_HD inline uint64_t bits_of(double v)
{
    uint64_t b;
    memcpy(&b, &v, sizeof b);
    return b;
}

constexpr double kEps = 1.1102230246251565e-16;                  // robust.hxx:281
constexpr double kErrOrient = (3.0 + 16.0 * kEps) * kEps;        // robust.hxx:282
constexpr double kErrIncircle = (10.0 + 96.0 * kEps) * kEps;     // robust.hxx:283
constexpr double kFilterFloor = 1e-250;                          // robust.hxx:289

QDT_HD inline void two_sum(double a, double b, double& x, double& y)
{
    x = a + b;
    const double bv = x - a;
    const double av = x - bv;
    y = (a - av) + (b - bv);
}

// a - b as two dyadic parts whose exact sum is the difference: the
// error-free (hi, lo) of two_sum, or (a, -b) itself when the rounded
// difference or one of two_sum's intermediate differences overflows
// (coordinates of magnitude near DBL_MAX; hi or lo is then infinite or NaN).
// The integer stages below only need hi + lo to be exact, not the parts to
// be non-overlapping.
QDT_HD inline void exact_diff(double a, double b, double& hi, double& lo)
{
    two_sum(a, -b, hi, lo);
    if (!(fabs(hi) <= 1.7976931348623157e308 && fabs(lo) <= 1.7976931348623157e308))
    {
        hi = a;
        lo = -b;
    }
}

// ---------------------------------------------------------------------------
// Integer decomposition of doubles
// ---------------------------------------------------------------------------

struct Dyad
{
    uint64_t m; // odd magnitude (0 for zero)
    int e;      // value = (neg ? -1 : 1) * m * 2^e
    int neg;
};

QDT_HD inline Dyad dyad_of(double v)
{
    Dyad d;
    const uint64_t b = bits_of(v);
    const int ex = (int)((b >> 52) & 0x7ff);
    uint64_t m = b & ((1ull << 52) - 1);
    int e;
    if (ex == 0)
        e = -1074;
    else
    {
        m |= 1ull << 52;
        e = ex - 1075;
    }
    d.neg = (int)(b >> 63);
    if (m == 0)
    {
        d.m = 0;
        d.e = 0;
        return d;
    }
    const int tz = __builtin_ctzll(m);
    d.m = m >> tz;
    d.e = e + tz;
    return d;
}

QDT_HD inline int bitlen64(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }

// Converts the components of k differences to integers in units of 2^E.
// Returns false when some difference needs more than 62 bits.
template <int K>
QDT_HD inline bool integerize(const double (&hi)[K], const double (&lo)[K], int64_t (&out)[K])
{
    Dyad dh[K], dl[K];
    int E = 1 << 20;
    bool any = false;
#pragma unroll
    for (int i = 0; i < K; ++i)
    {
        dh[i] = dyad_of(hi[i]);
        dl[i] = dyad_of(lo[i]);
        if (dh[i].m)
        {
            E = dh[i].e < E ? dh[i].e : E;
            any = true;
        }
        if (dl[i].m)
        {
            E = dl[i].e < E ? dl[i].e : E;
            any = true;
        }
    }
    if (!any)
    {
#pragma unroll
        for (int i = 0; i < K; ++i) out[i] = 0;
        return true;
    }
#pragma unroll
    for (int i = 0; i < K; ++i)
    {
        int64_t v = 0;
        if (dh[i].m)
        {
            const int sh = dh[i].e - E;
            if (sh > 61 || bitlen64(dh[i].m) + sh > 61) return false;
            const int64_t t = (int64_t)(dh[i].m << sh);
            v = dh[i].neg ? -t : t;
        }
        if (dl[i].m)
        {
            const int sh = dl[i].e - E;
            if (sh > 61 || bitlen64(dl[i].m) + sh > 61) return false;
            const int64_t t = (int64_t)(dl[i].m << sh);
            v += dl[i].neg ? -t : t;
        }
        // |hi|, |lo| < 2^61 so |v| < 2^62
        out[i] = v;
    }
    return true;
}

// ---------------------------------------------------------------------------
// 256-bit two's complement helpers
// ---------------------------------------------------------------------------

struct I256
{
    uint64_t w[4];
};

QDT_HD inline I256 umul128x128(unsigned __int128 a, unsigned __int128 b)
{
    const uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64);
    const uint64_t b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
    const unsigned __int128 p00 = (unsigned __int128)a0 * b0;
    const unsigned __int128 p01 = (unsigned __int128)a0 * b1;
    const unsigned __int128 p10 = (unsigned __int128)a1 * b0;
    const unsigned __int128 p11 = (unsigned __int128)a1 * b1;
    I256 r;
    r.w[0] = (uint64_t)p00;
    unsigned __int128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;
    r.w[1] = (uint64_t)mid;
    unsigned __int128 hi = (mid >> 64) + (p01 >> 64) + (p10 >> 64) + (uint64_t)p11;
    r.w[2] = (uint64_t)hi;
    r.w[3] = (uint64_t)(hi >> 64) + (uint64_t)(p11 >> 64);
    return r;
}

QDT_HD inline I256 neg256(I256 a)
{
    I256 r;
    uint64_t c = 1;
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        const uint64_t v = ~a.w[i] + c;
        c = (c && v == 0) ? 1 : 0;
        r.w[i] = v;
    }
    return r;
}

QDT_HD inline I256 add256(I256 a, I256 b)
{
    I256 r;
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        const uint64_t s = a.w[i] + b.w[i];
        const uint64_t c1 = s < a.w[i];
        const uint64_t s2 = s + c;
        const uint64_t c2 = s2 < s;
        r.w[i] = s2;
        c = c1 | c2;
    }
    return r;
}

QDT_HD inline int sign256(I256 a)
{
    if ((int64_t)a.w[3] < 0) return -1;
    return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) ? 1 : 0;
}

// lift (>= 0) times a signed 128-bit cross term
QDT_HD inline I256 mul_lift_cross(unsigned __int128 lift, __int128 cross)
{
    const int neg = cross < 0;
    const unsigned __int128 mag = neg ? (unsigned __int128)(-cross) : (unsigned __int128)cross;
    const I256 p = umul128x128(lift, mag);
    return neg ? neg256(p) : p;
}

// ---------------------------------------------------------------------------
// General exact path (any exponent spread).  Each difference is the sum of
// at most two dyads m * 2^e (m < 2^53), so the determinant expands into a sum
// of signed monomials, each a product of two (orient2d) or four (incircle)
// dyads: a <= 212-bit integer times a power of two.  The monomials are added
// into one two's-complement fixed-point accumulator whose unit is the
// smallest possible exponent.  Doubles span 2^-1074 .. 2^1024, so a degree-4
// monomial is shifted by at most 4 * 2100 bits: 136 64-bit limbs suffice.
// ---------------------------------------------------------------------------

constexpr int kAccLimbs = 136;

struct Acc
{
    uint64_t w[kAccLimbs];
};

// acc += (neg ? -v : v) * 2^shift, v a 256-bit magnitude
QDT_HD_NOINLINE void acc_add(Acc& acc, const I256& v, int shift, int neg)
{
    const int q = shift >> 6, r = shift & 63;
    uint64_t part[5];
    part[0] = v.w[0] << r;
    for (int i = 1; i < 4; ++i) part[i] = r ? (v.w[i] << r) | (v.w[i - 1] >> (64 - r)) : v.w[i];
    part[4] = r ? v.w[3] >> (64 - r) : 0;
    if (!neg)
    {
        uint64_t c = 0;
        for (int i = q; i < kAccLimbs; ++i)
        {
            const uint64_t a = i - q < 5 ? part[i - q] : 0;
            if (i - q >= 5 && c == 0) break;
            const uint64_t s1 = acc.w[i] + a;
            const uint64_t c1 = s1 < a;
            const uint64_t s2 = s1 + c;
            const uint64_t c2 = s2 < s1;
            acc.w[i] = s2;
            c = c1 | c2;
        }
    }
    else
    {
        uint64_t b = 0;
        for (int i = q; i < kAccLimbs; ++i)
        {
            const uint64_t a = i - q < 5 ? part[i - q] : 0;
            if (i - q >= 5 && b == 0) break;
            const uint64_t d1 = acc.w[i] - a;
            const uint64_t b1 = acc.w[i] < a;
            const uint64_t d2 = d1 - b;
            const uint64_t b2 = d1 < b;
            acc.w[i] = d2;
            b = b1 | b2;
        }
    }
}

QDT_HD inline int acc_sign(const Acc& acc)
{
    if ((int64_t)acc.w[kAccLimbs - 1] < 0) return -1;
    for (int i = 0; i < kAccLimbs; ++i)
        if (acc.w[i]) return 1;
    return 0;
}

// components of k differences as dyads: slot 2i = hi part, 2i+1 = lo part
template <int K>
QDT_HD inline int split_dyads(const double (&hi)[K], const double (&lo)[K], Dyad (&dy)[2 * K])
{
    int E = 1 << 20;
    for (int i = 0; i < K; ++i)
    {
        dy[2 * i] = dyad_of(hi[i]);
        dy[2 * i + 1] = dyad_of(lo[i]);
        if (dy[2 * i].m && dy[2 * i].e < E) E = dy[2 * i].e;
        if (dy[2 * i + 1].m && dy[2 * i + 1].e < E) E = dy[2 * i + 1].e;
    }
    return E;
}

QDT_HD_NOINLINE int orient_general(const double (&hi)[4], const double (&lo)[4])
{
    // order: acx, acy, bcx, bcy;  det = acx*bcy - acy*bcx
    Dyad dy[8];
    const int E = split_dyads<4>(hi, lo, dy);
    Acc acc;
    for (int i = 0; i < kAccLimbs; ++i) acc.w[i] = 0;
    const int terms[2][3] = { { 0, 3, 0 }, { 1, 2, 1 } }; // (u, v, negate)
    for (int t = 0; t < 2; ++t)
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
            {
                const Dyad& x = dy[2 * terms[t][0] + a];
                const Dyad& y = dy[2 * terms[t][1] + b];
                if (!x.m || !y.m) continue;
                I256 v = umul128x128((unsigned __int128)x.m, (unsigned __int128)y.m);
                const int sh = (x.e - E) + (y.e - E);
                acc_add(acc, v, sh, x.neg ^ y.neg ^ terms[t][2]);
            }
    return acc_sign(acc);
}

QDT_HD_NOINLINE int incircle_general(const double (&hi)[6], const double (&lo)[6])
{
    // order: adx, ady, bdx, bdy, cdx, cdy
    Dyad dy[12];
    const int E = split_dyads<6>(hi, lo, dy);
    Acc acc;
    for (int i = 0; i < kAccLimbs; ++i) acc.w[i] = 0;
    for (int k = 0; k < 3; ++k)
    {
        const int A = 2 * k, B = 2 * ((k + 1) % 3), C = 2 * ((k + 2) % 3);
        // lift = A.x^2 + A.y^2 ; cross = B.x*C.y - C.x*B.y
        const int lift_u[2] = { A, A + 1 };
        const int cross_u[2][3] = { { B, C + 1, 0 }, { C, B + 1, 1 } };
        for (int l = 0; l < 2; ++l)
            for (int c = 0; c < 2; ++c)
                for (int i1 = 0; i1 < 2; ++i1)
                    for (int i2 = 0; i2 < 2; ++i2)
                        for (int i3 = 0; i3 < 2; ++i3)
                            for (int i4 = 0; i4 < 2; ++i4)
                            {
                                const Dyad& p = dy[2 * lift_u[l] + i1];
                                const Dyad& q = dy[2 * lift_u[l] + i2];
                                const Dyad& r = dy[2 * cross_u[c][0] + i3];
                                const Dyad& t = dy[2 * cross_u[c][1] + i4];
                                if (!p.m || !q.m || !r.m || !t.m) continue;
                                const unsigned __int128 pq = (unsigned __int128)p.m * q.m;
                                const unsigned __int128 rt = (unsigned __int128)r.m * t.m;
                                const I256 v = umul128x128(pq, rt);
                                const int sh = (p.e - E) + (q.e - E) + (r.e - E) + (t.e - E);
                                acc_add(acc, v, sh, p.neg ^ q.neg ^ r.neg ^ t.neg ^ cross_u[c][2]);
                            }
    }
    return acc_sign(acc);
}

// ---------------------------------------------------------------------------
// Exact predicates
// ---------------------------------------------------------------------------

QDT_HD_NOINLINE int orient2d_exact(double ax, double ay, double bx, double by, double cx, double cy)
{
    double hi[4], lo[4];
    exact_diff(ax, cx, hi[0], lo[0]);
    exact_diff(ay, cy, hi[1], lo[1]);
    exact_diff(bx, cx, hi[2], lo[2]);
    exact_diff(by, cy, hi[3], lo[3]);
    int64_t d[4];
    if (integerize<4>(hi, lo, d))
    {
        const __int128 det = (__int128)d[0] * d[3] - (__int128)d[1] * d[2];
        return det > 0 ? 1 : (det < 0 ? -1 : 0);
    }
    return orient_general(hi, lo);
}

QDT_HD_NOINLINE int incircle_exact(double ax, double ay, double bx, double by, double cx, double cy,
                                           double dx, double dy)
{
    double hi[6], lo[6];
    exact_diff(ax, dx, hi[0], lo[0]);
    exact_diff(ay, dy, hi[1], lo[1]);
    exact_diff(bx, dx, hi[2], lo[2]);
    exact_diff(by, dy, hi[3], lo[3]);
    exact_diff(cx, dx, hi[4], lo[4]);
    exact_diff(cy, dy, hi[5], lo[5]);
    int64_t v[6];
    if (integerize<6>(hi, lo, v))
    {
        const __int128 adx = v[0], ady = v[1], bdx = v[2], bdy = v[3], cdx = v[4], cdy = v[5];
        const unsigned __int128 al = (unsigned __int128)(adx * adx) + (unsigned __int128)(ady * ady);
        const unsigned __int128 bl = (unsigned __int128)(bdx * bdx) + (unsigned __int128)(bdy * bdy);
        const unsigned __int128 cl = (unsigned __int128)(cdx * cdx) + (unsigned __int128)(cdy * cdy);
        const __int128 c1 = bdx * cdy - cdx * bdy;
        const __int128 c2 = cdx * ady - adx * cdy;
        const __int128 c3 = adx * bdy - bdx * ady;
        I256 s = mul_lift_cross(al, c1);
        s = add256(s, mul_lift_cross(bl, c2));
        s = add256(s, mul_lift_cross(cl, c3));
        return sign256(s);
    }
    return incircle_general(hi, lo);
}

// Filter stage: true when the FP64 determinant det has the sign of the exact
// one (robust.hxx:281-289 bounds).
QDT_HD inline bool orient2d_filter(double ax, double ay, double bx, double by, double cx, double cy, double& det)
{
    const double detleft = (ax - cx) * (by - cy);
    const double detright = (ay - cy) * (bx - cx);
    det = detleft - detright;
    const double mag = fabs(detleft) + fabs(detright);
    const double bound = kErrOrient * mag;
    return (det > bound || -det > bound) && mag > kFilterFloor;
}

QDT_HD inline bool incircle_filter(double ax, double ay, double bx, double by, double cx, double cy, double dx,
                                   double dy, double& det)
{
    const double adx = ax - dx, ady = ay - dy;
    const double bdx = bx - dx, bdy = by - dy;
    const double cdx = cx - dx, cdy = cy - dy;
    const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const double cdxady = cdx * ady, adxcdy = adx * cdy;
    const double adxbdy = adx * bdy, bdxady = bdx * ady;
    const double alift = adx * adx + ady * ady;
    const double blift = bdx * bdx + bdy * bdy;
    const double clift = cdx * cdx + cdy * cdy;
    det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
    const double permanent = (fabs(bdxcdy) + fabs(cdxbdy)) * alift + (fabs(cdxady) + fabs(adxcdy)) * blift +
        (fabs(adxbdy) + fabs(bdxady)) * clift;
    const double bound = kErrIncircle * permanent;
    return (det > bound || -det > bound) && permanent > kFilterFloor;
}

// The filters' determinants with a bound fixed in advance: valid for points
// whose coordinate differences are at most wx in x and wy in y in magnitude
// (a cell's extent).  The permanents are then at most 2 wx wy (orient2d) and
// 6 wx wy (wx^2 + wy^2) (incircle), so kErr * that bound is at least the
// run-time bound of the filters above; static_bounds returns both (0: not
// usable, the run-time filters apply) with a margin that covers the rounding
// of the computed differences, products and of the bound itself.
QDT_HD inline void static_bounds(double wx, double wy, double& b_orient, double& b_incircle)
{
    const double vx = wx * (1.0 + 0x1p-30), vy = wy * (1.0 + 0x1p-30);
    const double o = 2.0 * vx * vy, i = 6.0 * (vx * vy) * (vx * vx + vy * vy);
    const bool ok = o > kFilterFloor && i > kFilterFloor && i < 1e300;
    b_orient = ok ? kErrOrient * o * (1.0 + 0x1p-30) : 0.0;
    b_incircle = ok ? kErrIncircle * i * (1.0 + 0x1p-30) : 0.0;
}

QDT_HD inline double orient2d_det(double ax, double ay, double bx, double by, double cx, double cy)
{
    return (ax - cx) * (by - cy) - (ay - cy) * (bx - cx);
}

QDT_HD inline double incircle_det(double ax, double ay, double bx, double by, double cx, double cy, double dx,
                                  double dy)
{
    const double adx = ax - dx, ady = ay - dy;
    const double bdx = bx - dx, bdy = by - dy;
    const double cdx = cx - dx, cdy = cy - dy;
    const double alift = adx * adx + ady * ady;
    const double blift = bdx * bdx + bdy * bdy;
    const double clift = cdx * cdx + cdy * cdy;
    return alift * (bdx * cdy - cdx * bdy) + blift * (cdx * ady - adx * cdy) + clift * (adx * bdy - bdx * ady);
}

// Filtered predicates.  The return value's sign is the answer.  nexact counts
// the calls the filter could not decide.
QDT_HD inline double orient2d(double ax, double ay, double bx, double by, double cx, double cy,
                                           unsigned& nexact)
{
    double det;
    if (orient2d_filter(ax, ay, bx, by, cx, cy, det)) return det;
    ++nexact;
    return (double)orient2d_exact(ax, ay, bx, by, cx, cy);
}

QDT_HD inline double incircle(double ax, double ay, double bx, double by, double cx, double cy,
                                           double dx, double dy, unsigned& nexact)
{
    double det;
    if (incircle_filter(ax, ay, bx, by, cx, cy, dx, dy, det)) return det;
    ++nexact;
    return (double)incircle_exact(ax, ay, bx, by, cx, cy, dx, dy);
}

} // namespace qdt
