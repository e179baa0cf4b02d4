// Correctly rounded FP64 sin, cos, asin and atan2 for host and device code.
//
// The reference computes its geometry with .NET's Math.Sin / Cos / Asin / Atan2 (GeoMath.cs:49-80),
// which return whatever the platform's C library gives; those results are not correctly rounded (glibc
// 2.35 misrounds 0.1-0.6 % of latitudes and longitudes of a regional grid) and differ between libraries
// in their last bit. Where a ray step equals the piece length of RayTracing.Kernel (RayTracing.cs:139,
// a shared forward and inversion grid), ceil(len / h) is decided by that last bit, so a device path that
// is to reproduce the CPU path's cell lists must round exactly as it does. The one rounding every
// conforming platform can share is the correct one, so the device and the CPU restatement both use it.
//
// Method: Ziv's strategy (A. Ziv, "Fast evaluation of elementary mathematical functions with correctly
// rounded last bit", ACM TOMS 17(3), 1991), the quick phase only: each function is evaluated in
// double-double arithmetic (Dekker's products and Knuth's sums) with a relative error below 2^-100, and
// the double nearest to the double-double value is returned. Ziv's rounding test (ok below) says
// whether that double is provably the correctly rounded value; it fails with probability ~2^-46 per
// call, and then the quick phase's double is returned anyway (correct unless the exact value lies
// within 2^-100 of a midpoint between doubles). Pieces:
//   sin / cos: Cody and Waite's reduction by pi/2 in four parts (|x| <= 1024), a table of sin and cos
//     at i/64 (Tang's table-driven method), degree-11 / degree-12 Taylor polynomials of the remainder
//     (|t| <= 1/128), the leading terms in double-double.
//   atan2: the octant reduction to atan(q), q in [0, 1], a table of atan(i/64) and the series of
//     atan((q - c) / (1 + q c)) to degree 15.
//   asin(z) = atan2(z, sqrt((1 - z)(1 + z))), the root in double-double.
// Arguments outside these ranges (|x| > 1024 for sin / cos, non-finite, zero or extreme-ratio atan2
// operands) fall back to the library's functions, whose results there are exact special values or
// are not used by the geometry.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "crmath_tables.hpp"

namespace ts {
namespace cr {

struct dd {
    double hi, lo;
};

// a + b = s + e exactly (Knuth)
__host__ __device__ inline dd two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
// a + b = s + e exactly when |a| >= |b| (Dekker)
__host__ __device__ inline dd fast_two_sum(double a, double b)
{
    const double s = a + b;
    return {s, b - (s - a)};
}
// a * b = p + e exactly (the fused multiply-add gives the rounding error)
__host__ __device__ inline dd two_prod(double a, double b)
{
    const double p = a * b;
    return {p, fma(a, b, -p)};
}
__host__ __device__ inline dd dd_add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
__host__ __device__ inline dd dd_add_d(dd a, double b)
{
    dd s = two_sum(a.hi, b);
    s.lo += a.lo;
    return fast_two_sum(s.hi, s.lo);
}
__host__ __device__ inline dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
__host__ __device__ inline dd dd_mul(dd a, dd b)
{
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo);
}
__host__ __device__ inline dd dd_mul_d(dd a, double b)
{
    dd p = two_prod(a.hi, b);
    p.lo += a.lo * b;
    return fast_two_sum(p.hi, p.lo);
}
// a / b: three quotient digits, each from the exact remainder of the previous ones
__host__ __device__ inline dd dd_div(dd a, dd b)
{
    const double q1 = a.hi / b.hi;
    dd r = dd_add(a, dd_neg(dd_mul_d(b, q1)));
    const double q2 = r.hi / b.hi;
    r = dd_add(r, dd_neg(dd_mul_d(b, q2)));
    const double q3 = r.hi / b.hi;
    dd q = fast_two_sum(q1, q2);
    return dd_add_d(q, q3);
}
// sqrt(a) for a > 0: the double root and one correction from its exact residual
__host__ __device__ inline dd dd_sqrt(dd a)
{
    const double s = sqrt(a.hi);
    const double e = fma(-s, s, a.hi) + a.lo;
    return fast_two_sum(s, e / (2 * s));
}

// The double nearest to y (normalised: y.hi = RN(y.hi + y.lo)), and Ziv's test that every value within
// a relative distance eps of y rounds to it.
__host__ __device__ inline double round_dd(dd y, double eps, bool *ok)
{
    if (ok) {
        const double e = eps * fabs(y.hi);
        *ok = (y.hi + (y.lo + e)) == (y.hi + (y.lo - e));
    }
    return y.hi + y.lo;
}
constexpr double kEps = 0x1p-98;  // bound used in the test: four times the evaluations' 2^-100

// sin and cos of r (|r| <= pi/4 + 2^-40) in double-double
__host__ __device__ inline void sincos_reduced(dd r, dd &s, dd &c)
{
    const bool neg = r.hi < 0;
    if (neg) r = dd_neg(r);
    int i = (int)rint(r.hi * 64);
    i = i < 0 ? 0 : (i > 51 ? 51 : i);  // |r| <= pi/4 + 2^-40 keeps i <= 50; the clamp keeps reads in the table
    // t = r - i/64: the difference of the heads is exact (Sterbenz), the tail is added exactly
    const dd t = i == 0 ? r : two_sum(r.hi - i * 0x1p-6, r.lo);
    const dd t2 = dd_mul(t, t);
    // sin t = t + t t^2 (S3 + t^2 (S5 + t^2 (S7 + t^2 (S9 + t^2 S11))))
    const double rs2 = kS7 + t2.hi * (kS9 + t2.hi * kS11);
    const dd rs1 = dd_add_d(dd{kS5_hi, kS5_lo}, t2.hi * rs2);
    const dd qs = dd_add(dd{kS3_hi, kS3_lo}, dd_mul(t2, rs1));
    const dd st = dd_add(t, dd_mul(t, dd_mul(t2, qs)));
    // cos t = 1 + t^2 (-1/2 + t^2 (C4 + t^2 (C6 + t^2 (C8 + t^2 (C10 + t^2 C12)))))
    const double uc3 = kC8 + t2.hi * (kC10 + t2.hi * kC12);
    const dd uc2 = dd_add_d(dd{kC6_hi, kC6_lo}, t2.hi * uc3);
    const dd uc1 = dd_add(dd{kC4_hi, kC4_lo}, dd_mul(t2, uc2));
    const dd pc = dd_add_d(dd_mul(t2, uc1), -0.5);
    const dd ct = dd_add_d(dd_mul(t2, pc), 1.0);
    if (i == 0) {
        s = st;
        c = ct;
    } else {
        // sin(a + t) = sin a cos t + cos a sin t ; cos(a + t) = cos a cos t - sin a sin t
        const double *e = kSinCos64[i];
        const dd sa{e[0], e[1]}, ca{e[2], e[3]};
        s = dd_add(dd_mul(sa, ct), dd_mul(ca, st));
        c = dd_add(dd_mul(ca, ct), dd_neg(dd_mul(sa, st)));
    }
    if (neg) s = dd_neg(s);
}

// x = k pi/2 + r with |r| <= pi/4 (+ rounding), r in double-double, for |x| <= 1024; returns k mod 4
__host__ __device__ inline int reduce_pio2(double x, dd &r)
{
    if (fabs(x) <= 0x1.921fb54442d18p-1) {
        r = {x, 0};
        return 0;
    }
    const double k = rint(x * kTwoOverPi);
    // x - k p1 is exact (Sterbenz), k p2 and k p3's split product are exact
    dd a = two_sum(x - k * kPio2_1, -k * kPio2_2);
    a = dd_add(a, dd_neg(two_prod(k, kPio2_3)));
    r = dd_add_d(a, -k * kPio2_4);
    return (int)((long long)k & 3);
}

__host__ __device__ inline void sincos_dd(double x, dd &s, dd &c)
{
    dd r;
    const int q = reduce_pio2(x, r);
    dd sr, cr;
    sincos_reduced(r, sr, cr);
    switch (q) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = dd_neg(sr); break;
    case 2: s = dd_neg(sr); c = dd_neg(cr); break;
    default: s = dd_neg(cr); c = sr; break;
    }
}

__host__ __device__ inline bool sincos_in_range(double x) { return fabs(x) <= 1024.0; }

__host__ __device__ inline double sin(double x, bool *ok = nullptr)
{
    if (ok) *ok = true;
    if (!sincos_in_range(x)) return ::sin(x);
    if (fabs(x) < 0x1p-27) return x;  // |sin x - x| < 2^-56.6 |x|: x is the nearest double
    dd s, c;
    sincos_dd(x, s, c);
    return round_dd(s, kEps, ok);
}

__host__ __device__ inline double cos(double x, bool *ok = nullptr)
{
    if (ok) *ok = true;
    if (!sincos_in_range(x)) return ::cos(x);
    if (fabs(x) < 0x1p-27) return 1.0;  // 1 - cos x < 2^-55: 1 is the nearest double
    dd s, c;
    sincos_dd(x, s, c);
    return round_dd(c, kEps, ok);
}

// both at once (one reduction and one polynomial pair)
__host__ __device__ inline void sincos(double x, double &s, double &c)
{
    if (!sincos_in_range(x)) {
        s = ::sin(x);
        c = ::cos(x);
        return;
    }
    if (fabs(x) < 0x1p-27) {
        s = x;
        c = 1.0;
        return;
    }
    dd sd, cd;
    sincos_dd(x, sd, cd);
    s = round_dd(sd, kEps, nullptr);
    c = round_dd(cd, kEps, nullptr);
}

// atan(q) for q in [0, 1 + 2^-50] in double-double
__host__ __device__ inline dd atan01(dd q)
{
    int i = (int)rint(q.hi * 64);
    i = i < 0 ? 0 : (i > 64 ? 64 : i);  // q in [0, 1 + 2^-50] keeps i in range; the clamp keeps reads in the table
    dd u;
    if (i == 0) {
        u = q;
    } else {
        const double c = i * 0x1p-6;
        // (q - c) / (1 + q c): the difference of the heads is exact (Sterbenz)
        const dd num = two_sum(q.hi - c, q.lo);
        const dd den = dd_add_d(dd_mul_d(q, c), 1.0);
        u = dd_div(num, den);
    }
    const dd u2 = dd_mul(u, u);
    // atan u = u + u u^2 (A3 + u^2 (A5 + u^2 (A7 + u^2 (A9 + u^2 (A11 + u^2 (A13 + u^2 A15))))))
    const double w3 = kA9 + u2.hi * (kA11 + u2.hi * (kA13 + u2.hi * kA15));
    const dd w2 = dd_add_d(dd{kA7_hi, kA7_lo}, u2.hi * w3);
    const dd w1 = dd_add(dd{kA5_hi, kA5_lo}, dd_mul(u2, w2));
    const dd v = dd_add(dd{kA3_hi, kA3_lo}, dd_mul(u2, w1));
    dd at = dd_add(u, dd_mul(u, dd_mul(u2, v)));
    if (i != 0) at = dd_add(dd{kAtan64[i][0], kAtan64[i][1]}, at);
    return at;
}

// atan2 of two non-negative double-double operands (not both zero), in [0, pi/2]
__host__ __device__ inline dd atan2_pos(dd y, dd x)
{
    if (y.hi <= x.hi) return atan01(dd_div(y, x));
    return dd_add(dd{kPio2_hi, kPio2_lo}, dd_neg(atan01(dd_div(x, y))));
}

__host__ __device__ inline double atan2(double y, double x, bool *ok = nullptr)
{
    if (ok) *ok = true;
    const double ay = fabs(y), ax = fabs(x);
    // zeros, infinities, NaN: exact special values from the library
    if (!(ay > 0) || !(ax > 0) || ay == INFINITY || ax == INFINITY) return ::atan2(y, x);
    // ratios far from 1 (beyond 2^900) are outside the double-double division's exponent range
    const int ey = ilogb(ay), ex = ilogb(ax);
    if (ey - ex > 900 || ex - ey > 900) return ::atan2(y, x);
    dd base = atan2_pos(dd{ay, 0}, dd{ax, 0});
    if (x < 0) base = dd_add(dd{kPi_hi, kPi_lo}, dd_neg(base));
    const double r = round_dd(base, kEps, ok);
    return y < 0 ? -r : r;
}

__host__ __device__ inline double asin(double z, bool *ok = nullptr)
{
    if (ok) *ok = true;
    const double az = fabs(z);
    if (!(az <= 1.0)) return ::asin(z);       // NaN, |z| > 1
    if (az < 0x1p-28) return z;               // asin z - z < 2^-58 |z|: z is the nearest double
    dd base;
    if (az == 1.0) {
        base = dd{kPio2_hi, kPio2_lo};
    } else {
        // 1 - z^2 = (1 - |z|)(1 + |z|) (the first factor exact for |z| >= 1/2), else 1 - z z
        dd d;
        if (az >= 0.5) d = dd_mul_d(two_sum(1.0, az), 1.0 - az);
        else d = dd_add(dd{1.0, 0}, dd_neg(two_prod(az, az)));
        base = atan2_pos(dd{az, 0}, dd_sqrt(d));
    }
    const double r = round_dd(base, kEps, ok);
    return z < 0 ? -r : r;
}

}  // namespace cr
}  // namespace ts
