// f32 transcendental functions of the detector, evaluated in double with only
// +, -, *, /, sqrt and exact scaling, then rounded once to f32: the correctly
// rounded f32 value unless the true value lies within ~1e-16 (relative) of a
// rounding midpoint. The operations and their order are fixed (no FMA
// contraction), so host and device produce the same bits.
//   exp:     x = k ln2 + r (Cody-Waite, ln2 in two pieces), Taylor to r^13
//   log:     x = m 2^e, m in [sqrt(1/2), sqrt(2)), 2 atanh((m-1)/(m+1)) to s^27
//   atan2:   quadrant reduction, atan t = pi/4 + atan((t-1)/(t+1)) above
//            tan(pi/8), one half-angle step, series to u^25
//   sin/cos: a = k pi/2 + r (three 33-bit pieces of pi/2), Taylor to r^23
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

namespace skh {
namespace fm {

constexpr double kPi = 3.14159265358979311600e+00;
constexpr double kPio2 = 1.57079632679489655800e+00;
constexpr double kPio4 = 7.85398163397448278999e-01;
constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 32 trailing zero bits
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kPio2a = 1.57079632673412561417e+00;  // 33-bit pieces of pi/2
constexpr double kPio2b = 6.07710050650619224932e-11;
constexpr double kPio2c = 2.02226624879595063154e-21;

__device__ inline double xexp(double x) {
    if (x != x) return x;
    if (x > 709.0) return __builtin_inf();
    if (x < -745.0) return 0.0;
    const double k = rint(x * 1.44269504088896338700e+00);
    const double r = (x - k * kLn2Hi) - k * kLn2Lo;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int)k);
}

__device__ inline double xlog(double x) {
    int e;
    double m = frexp(x, &e);
    if (m < 7.07106781186547572737e-01) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double p = 1.0 / 27.0;
    p = p * s2 + 1.0 / 25.0;
    p = p * s2 + 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    const double de = (double)e;
    return de * kLn2Hi + (de * kLn2Lo + 2.0 * s * p);
}

__device__ inline double xatan01(double t) {
    double base = 0.0;
    if (t > 4.14213562373095034466e-01) {
        t = (t - 1.0) / (t + 1.0);
        base = kPio4;
    }
    const double u = t / (1.0 + sqrt(1.0 + t * t));
    const double u2 = u * u;
    double p = 1.0 / 25.0;
    p = -(p * u2) + 1.0 / 23.0;
    p = -(p * u2) + 1.0 / 21.0;
    p = -(p * u2) + 1.0 / 19.0;
    p = -(p * u2) + 1.0 / 17.0;
    p = -(p * u2) + 1.0 / 15.0;
    p = -(p * u2) + 1.0 / 13.0;
    p = -(p * u2) + 1.0 / 11.0;
    p = -(p * u2) + 1.0 / 9.0;
    p = -(p * u2) + 1.0 / 7.0;
    p = -(p * u2) + 1.0 / 5.0;
    p = -(p * u2) + 1.0 / 3.0;
    p = -(p * u2) + 1.0;
    return base + 2.0 * (u * p);
}

__device__ inline double xatan2(double y, double x) {
    if (x != x || y != y) return x + y;
    const double ax = fabs(x), ay = fabs(y);
    double r;
    if (ay == 0.0) {
        r = signbit(x) ? kPi : 0.0;
    } else if (isinf(ax) && isinf(ay)) {
        r = signbit(x) ? 3.0 * kPio4 : kPio4;
    } else if (isinf(ay) || ax == 0.0) {
        r = kPio2;
    } else if (isinf(ax)) {
        r = signbit(x) ? kPi : 0.0;
    } else {
        r = ay <= ax ? xatan01(ay / ax) : kPio2 - xatan01(ax / ay);
        if (signbit(x)) r = kPi - r;
    }
    return signbit(y) ? -r : r;
}

__device__ inline void xsincos(double a, double* s_out, double* c_out) {
    if (!isfinite(a) || fabs(a) > 1e6) {
        *s_out = __builtin_nan("");
        *c_out = __builtin_nan("");
        return;
    }
    const double k = rint(a * 6.36619772367581382433e-01);
    const double r = ((a - k * kPio2a) - k * kPio2b) - k * kPio2c;
    const double r2 = r * r;
    double s = -1.0 / 25852016738884976640000.0;
    s = s * r2 + 1.0 / 51090942171709440000.0;
    s = s * r2 - 1.0 / 121645100408832000.0;
    s = s * r2 + 1.0 / 355687428096000.0;
    s = s * r2 - 1.0 / 1307674368000.0;
    s = s * r2 + 1.0 / 6227020800.0;
    s = s * r2 - 1.0 / 39916800.0;
    s = s * r2 + 1.0 / 362880.0;
    s = s * r2 - 1.0 / 5040.0;
    s = s * r2 + 1.0 / 120.0;
    s = s * r2 - 1.0 / 6.0;
    s = s * r2 + 1.0;
    s = s * r;
    double c = 1.0 / 1124000727777607680000.0;
    c = c * r2 - 1.0 / 2432902008176640000.0;
    c = c * r2 + 1.0 / 6402373705728000.0;
    c = c * r2 - 1.0 / 20922789888000.0;
    c = c * r2 + 1.0 / 87178291200.0;
    c = c * r2 - 1.0 / 479001600.0;
    c = c * r2 + 1.0 / 3628800.0;
    c = c * r2 - 1.0 / 40320.0;
    c = c * r2 + 1.0 / 720.0;
    c = c * r2 - 1.0 / 24.0;
    c = c * r2 + 0.5;
    c = 1.0 - c * r2;
    const int q = (int)(((long long)k) & 3);
    if (q == 0) {
        *s_out = s;
        *c_out = c;
    } else if (q == 1) {
        *s_out = c;
        *c_out = -s;
    } else if (q == 2) {
        *s_out = -s;
        *c_out = -c;
    } else {
        *s_out = -c;
        *c_out = s;
    }
}

__device__ inline float expf_cr(float x) { return (float)xexp((double)x); }
__device__ inline float atan2f_cr(float y, float x) { return (float)xatan2((double)y, (double)x); }
__device__ inline void sincosf_cr(float a, float* s, float* c) {
    double sd, cd;
    xsincos((double)a, &sd, &cd);
    *s = (float)sd;
    *c = (float)cd;
}
// b^y for finite b > 0 and finite y (the caller guarantees both).
__device__ inline float powf_pos(float b, float y) {
    return (float)xexp((double)y * xlog((double)b));
}
// |(dx, dy)| with exact squares in double, then to f32.
__device__ inline float hypotf_cr(float dx, float dy) {
    const double a = (double)dx, b = (double)dy;
    return (float)sqrt(a * a + b * b);
}

}  // namespace fm
}  // namespace skh
