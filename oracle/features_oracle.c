/*
 * CPU oracle of the feature detector (crates/core/src/features.rs:55-781):
 * Gaussian pyramid, DoG extrema with sub-pixel/sub-scale refinement, edge
 * test, dominant orientations and the 128-d descriptor. TEST INFRASTRUCTURE
 * ONLY (see oracle.h).
 *
 * Everything is f32 in the reference's operation order (-ffp-contract=off).
 * Two kinds of f32 transcendental calls occur:
 *  - on values the host computes once per image (Gaussian taps, the scale
 *    step 2^(1/s)): glibc's expf / powf, as the reference's Rust f32::exp /
 *    powf call them;
 *  - per pixel or per keypoint (atan2 of the gradient table, the orientation
 *    and descriptor weights, sin/cos of the keypoint angle, the keypoint
 *    sigma): orc_expf / orc_atan2f / orc_sincosf / orc_powf below. They
 *    evaluate the function in double precision with only +, -, *, /, sqrt
 *    and exact scaling (Cody-Waite range reduction, Taylor / atanh series)
 *    and round once to f32, so the result is the correctly rounded f32 value
 *    unless the true value lies within ~1e-16 (relative) of a rounding
 *    midpoint. glibc's expf/atan2f/sinf/cosf/powf agree with them except
 *    where glibc itself misrounds (tests/test_oracle_features.py measures
 *    this on millions of inputs). The device evaluates the same operations
 *    in the same order, so oracle and device agree bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---- f32 transcendentals evaluated in double --------------------------- */

static const double kPi = 3.14159265358979311600e+00;
static const double kPio2 = 1.57079632679489655800e+00;
static const double kPio4 = 7.85398163397448278999e-01;
/* ln 2 = kLn2Hi + kLn2Lo, kLn2Hi with 32 trailing zero bits (k * kLn2Hi exact) */
static const double kLn2Hi = 6.93147180369123816490e-01;
static const double kLn2Lo = 1.90821492927058770002e-10;
/* pi/2 in three pieces of 33 bits (k * piece exact for |k| < 2^20) */
static const double kPio2a = 1.57079632673412561417e+00;
static const double kPio2b = 6.07710050650619224932e-11;
static const double kPio2c = 2.02226624879595063154e-21;

/* e^x: x = k ln2 + r, |r| <= ln2/2, Taylor polynomial of degree 13 in r,
 * then scaled by 2^k. */
double orc_xexp(double x) {
    if (x != x) return x;
    if (x > 709.0) return INFINITY;
    if (x < -745.0) return 0.0;
    const double k = rint(x * 1.44269504088896338700e+00);
    const double r = (x - k * kLn2Hi) - k * kLn2Lo;
    double p = 1.0 / 6227020800.0; /* 1/13! */
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

/* ln x for finite x > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s), s = (m - 1)/(m + 1), series to s^27. */
double orc_xlog(double x) {
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

/* atan t for t in [0, 1]: t > tan(pi/8) -> pi/4 + atan((t-1)/(t+1)); then
 * the half-angle step atan t = 2 atan(t / (1 + sqrt(1 + t^2))) leaves
 * |u| <= 0.199 for a series to u^25. */
static double xatan01(double t) {
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

/* atan2(y, x) with IEEE sign conventions for zeros and infinities. */
double orc_xatan2(double y, double x) {
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

/* sin and cos: a = k pi/2 + r (three-piece Cody-Waite, |k| < 2^20), Taylor
 * series of sin r to r^23 and cos r to r^22, quadrant k mod 4. */
void orc_xsincos(double a, double* s_out, double* c_out) {
    if (!isfinite(a) || fabs(a) > 1e6) {
        *s_out = NAN;
        *c_out = NAN;
        return;
    }
    const double k = rint(a * 6.36619772367581382433e-01);
    const double r = ((a - k * kPio2a) - k * kPio2b) - k * kPio2c;
    const double r2 = r * r;
    double s = -1.0 / 25852016738884976640000.0; /* -1/23! */
    s = s * r2 + 1.0 / 51090942171709440000.0;   /* 1/21! */
    s = s * r2 - 1.0 / 121645100408832000.0;     /* 1/19! */
    s = s * r2 + 1.0 / 355687428096000.0;        /* 1/17! */
    s = s * r2 - 1.0 / 1307674368000.0;          /* 1/15! */
    s = s * r2 + 1.0 / 6227020800.0;             /* 1/13! */
    s = s * r2 - 1.0 / 39916800.0;               /* 1/11! */
    s = s * r2 + 1.0 / 362880.0;                 /* 1/9! */
    s = s * r2 - 1.0 / 5040.0;                   /* 1/7! */
    s = s * r2 + 1.0 / 120.0;                    /* 1/5! */
    s = s * r2 - 1.0 / 6.0;                      /* 1/3! */
    s = s * r2 + 1.0;
    s = s * r;
    double c = 1.0 / 1124000727777607680000.0; /* 1/22! */
    c = c * r2 - 1.0 / 2432902008176640000.0;  /* 1/20! */
    c = c * r2 + 1.0 / 6402373705728000.0;     /* 1/18! */
    c = c * r2 - 1.0 / 20922789888000.0;       /* 1/16! */
    c = c * r2 + 1.0 / 87178291200.0;          /* 1/14! */
    c = c * r2 - 1.0 / 479001600.0;            /* 1/12! */
    c = c * r2 + 1.0 / 3628800.0;              /* 1/10! */
    c = c * r2 - 1.0 / 40320.0;                /* 1/8! */
    c = c * r2 + 1.0 / 720.0;                  /* 1/6! */
    c = c * r2 - 1.0 / 24.0;                   /* 1/4! */
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

float orc_expf(float x) { return (float)orc_xexp((double)x); }
float orc_atan2f(float y, float x) { return (float)orc_xatan2((double)y, (double)x); }
void orc_sincosf(float a, float* s, float* c) {
    double sd, cd;
    orc_xsincos((double)a, &sd, &cd);
    *s = (float)sd;
    *c = (float)cd;
}
/* b^y for finite b > 0 (the detector's scale step; the reference calls powf
 * on 2^(1/s) with a refined layer coordinate). */
float orc_powf(float b, float y) {
    if (!(b > 0.0f) || !isfinite(b) || !isfinite(y)) return powf(b, y);
    return (float)orc_xexp((double)y * orc_xlog((double)b));
}
/* |(dx, dy)|: both squares are exact in double, one rounding for the sum and
 * one for the square root, then to f32. */
static float hypot_f(float dx, float dy) {
    const double a = (double)dx, b = (double)dy;
    return (float)sqrt(a * a + b * b);
}

/* ---- images -------------------------------------------------------------- */

static const float kTau = 6.28318548202514648438f; /* f32 TAU */
static const float kSqrt2 = 1.41421353816986083984f; /* f32 SQRT_2 */

/* f32::rem_euclid (Rust core): r = a % b; r < 0 ? r + |b| : r. */
static float rem_euclid(float a, float b) {
    const float r = fmodf(a, b);
    return r < 0.0f ? r + fabsf(b) : r;
}
/* f32 as isize / usize: saturating, NaN -> 0 (only small values occur). */
static long to_isize(float v) {
    if (v != v) return 0;
    if (v >= 9.2e18f) return 9223372036854775807L;
    if (v <= -9.2e18f) return (-9223372036854775807L - 1);
    return (long)v;
}
/* f32::powi (LLVM powi / compiler-rt __powisf2): binary exponentiation. */
static float powi_f(float a, int b) {
    const int recip = b < 0;
    float r = 1.0f;
    for (;;) {
        if (b & 1) r = r * a;
        b /= 2;
        if (b == 0) break;
        a = a * a;
    }
    return recip ? 1.0f / r : r;
}

/* GrayImage::upsample2 (features.rs:55-81). out: (2w) x (2h). */
void orc_upsample2(const float* img, size_t w, size_t h, float* out) {
    const size_t W = w * 2, H = h * 2;
    const long mw = (long)w - 1, mh = (long)h - 1;
    for (size_t y = 0; y < H; ++y) {
        const float sy = ((float)y + 0.5f) / 2.0f - 0.5f;
        const float y0 = floorf(sy);
        const float fy = sy - y0;
        long ya = to_isize(y0), yb = to_isize(y0) + 1;
        ya = ya < 0 ? 0 : (ya > mh ? mh : ya);
        yb = yb < 0 ? 0 : (yb > mh ? mh : yb);
        for (size_t x = 0; x < W; ++x) {
            const float sx = ((float)x + 0.5f) / 2.0f - 0.5f;
            const float x0 = floorf(sx);
            const float fx = sx - x0;
            long xa = to_isize(x0), xb = to_isize(x0) + 1;
            xa = xa < 0 ? 0 : (xa > mw ? mw : xa);
            xb = xb < 0 ? 0 : (xb > mw ? mw : xb);
            const float top = img[ya * w + xa] * (1.0f - fx) + img[ya * w + xb] * fx;
            const float bot = img[yb * w + xa] * (1.0f - fx) + img[yb * w + xb] * fx;
            out[y * W + x] = top * (1.0f - fy) + bot * fy;
        }
    }
}

/* GrayImage::downsample (features.rs:200-210): out (w/2) x (h/2). */
void orc_downsample(const float* img, size_t w, size_t h, float* out) {
    const size_t W = w / 2, H = h / 2;
    for (size_t y = 0; y < H; ++y)
        for (size_t x = 0; x < W; ++x) out[y * W + x] = img[(2 * y) * w + 2 * x];
}

/* Taps of gaussian_blur (features.rs:215-221): returns r, writes 2r+1 taps
 * (cap entries available; returns -1 if cap is short). */
long orc_gaussian_taps(float sigma, float* k, size_t cap) {
    float rv = ceilf(3.0f * sigma);
    rv = rv > 1.0f ? rv : 1.0f; /* f32::max: NaN picks the other operand */
    const long r = to_isize(rv);
    if ((size_t)(2 * r + 1) > cap) return -1;
    float s = -0.0f;
    for (long i = -r; i <= r; ++i) {
        k[i + r] = expf(-(float)(i * i) / (2.0f * sigma * sigma));
    }
    for (long i = 0; i <= 2 * r; ++i) s = s + k[i];
    for (long i = 0; i <= 2 * r; ++i) k[i] = k[i] / s;
    return r;
}

/* gaussian_blur (features.rs:214-256): horizontal pass over the edge-clamped
 * row, then the vertical pass accumulating rows in tap order. */
int orc_gaussian_blur(const float* img, size_t w, size_t h, float sigma, float* out) {
    float k[1025];
    const long r = orc_gaussian_taps(sigma, k, 1025);
    if (r < 0) return -1;
    float* tmp = (float*)malloc(sizeof(float) * (w * h > 0 ? w * h : 1));
    if (!tmp) return -1;
    const long W = (long)w, H = (long)h;
    for (long y = 0; y < H; ++y) {
        const float* src = img + y * W;
        for (long x = 0; x < W; ++x) {
            float acc = 0.0f;
            for (long t = 0; t <= 2 * r; ++t) {
                long xx = x + t - r;
                xx = xx < 0 ? 0 : (xx > W - 1 ? W - 1 : xx);
                acc = acc + k[t] * src[xx];
            }
            tmp[y * W + x] = acc;
        }
    }
    for (long y = 0; y < H; ++y) {
        float* row = out + y * W;
        for (long x = 0; x < W; ++x) row[x] = 0.0f;
        for (long j = 0; j <= 2 * r; ++j) {
            long yy = y + j - r;
            yy = yy < 0 ? 0 : (yy > H - 1 ? H - 1 : yy);
            const float* src = tmp + yy * W;
            for (long x = 0; x < W; ++x) row[x] = row[x] + k[j] * src[x];
        }
    }
    free(tmp);
    return 0;
}

/* pixel_grad (features.rs:325-331): central differences, magnitude and
 * atan2(gy, gx). */
static void pixel_grad(const float* img, size_t w, size_t ux, size_t uy, float* mag, float* at) {
    const float gx = img[uy * w + ux + 1] - img[uy * w + ux - 1];
    const float gy = img[(uy + 1) * w + ux] - img[(uy - 1) * w + ux];
    *mag = sqrtf(gx * gx + gy * gy);
    *at = orc_atan2f(gy, gx);
}

/* orientations_with (features.rs:361-413). grad(ux, uy) is the gradient
 * table: mag/ang arrays of a w x h layer. Returns the angle count. */
static size_t orientations_with(size_t width, size_t height, const float* gmag,
                                const float* gang, float x, float y, float sigma, float* out) {
    enum { NB = 36 };
    const float sw = 1.5f * sigma;
    const long r = to_isize(roundf(3.0f * sw));
    const long xi = to_isize(roundf(x)), yi = to_isize(roundf(y));
    const long w = (long)width, h = (long)height;
    float hist[NB];
    for (int i = 0; i < NB; ++i) hist[i] = 0.0f;
    for (long dy = -r; dy <= r; ++dy) {
        for (long dx = -r; dx <= r; ++dx) {
            const long px = xi + dx, py = yi + dy;
            if (px < 1 || py < 1 || px >= w - 1 || py >= h - 1) continue;
            const float mag = gmag[py * w + px], at = gang[py * w + px];
            const float fx = (float)px - x, fy = (float)py - y;
            const float wgt = orc_expf(-(fx * fx + fy * fy) / (2.0f * sw * sw));
            const float ang = rem_euclid(at, kTau);
            const size_t bin = (size_t)to_isize(roundf(ang / kTau * (float)NB)) % NB;
            hist[bin] = hist[bin] + wgt * mag;
        }
    }
    for (int pass = 0; pass < 6; ++pass) {
        float prev[NB];
        memcpy(prev, hist, sizeof prev);
        for (int i = 0; i < NB; ++i)
            hist[i] = (prev[(i + NB - 1) % NB] + prev[i] + prev[(i + 1) % NB]) / 3.0f;
    }
    float mx = 0.0f;
    for (int i = 0; i < NB; ++i) mx = fmaxf(mx, hist[i]);
    if (mx <= 0.0f) return 0;
    size_t n = 0;
    for (int i = 0; i < NB; ++i) {
        const float l = hist[(i + NB - 1) % NB], c = hist[i], rr = hist[(i + 1) % NB];
        if (c >= 0.8f * mx && c > l && c > rr) {
            const float off = 0.5f * (l - rr) / (l - 2.0f * c + rr);
            out[n++] = rem_euclid(((float)i + off) / (float)NB * kTau, kTau);
        }
    }
    return n;
}

/* describe_with (features.rs:521-592). */
static void describe_with(size_t width, size_t height, const float* gmag, const float* gang,
                          float x, float y, float sigma, float angle, float* hist) {
    enum { NC = 4, NO = 8 };
    const float cell = 3.0f * sigma;
    const long r = to_isize(roundf(cell * kSqrt2 * ((float)NC + 1.0f) * 0.5f));
    float c, sn;
    orc_sincosf(angle, &sn, &c);
    const long xi = to_isize(roundf(x)), yi = to_isize(roundf(y));
    const long w = (long)width, h = (long)height;
    for (int i = 0; i < ORC_DESC_LEN; ++i) hist[i] = 0.0f;
    for (long dy = -r; dy <= r; ++dy) {
        for (long dx = -r; dx <= r; ++dx) {
            const long px = xi + dx, py = yi + dy;
            if (px < 1 || py < 1 || px >= w - 1 || py >= h - 1) continue;
            const float fx = (float)px - x, fy = (float)py - y;
            const float u = (c * fx + sn * fy) / cell;
            const float v = (-sn * fx + c * fy) / cell;
            const float bu = u + (float)NC / 2.0f - 0.5f, bv = v + (float)NC / 2.0f - 0.5f;
            if (bu <= -1.0f || bv <= -1.0f || bu >= (float)NC || bv >= (float)NC) continue;
            const float mag = gmag[py * w + px], at = gang[py * w + px];
            const float ori = rem_euclid(at - angle, kTau);
            const float bo = ori / kTau * (float)NO;
            const float wgt = orc_expf(-(u * u + v * v) / (2.0f * powi_f((float)NC / 2.0f, 2)));
            const float m = mag * wgt;
            const float u0 = floorf(bu), v0 = floorf(bv), o0 = floorf(bo);
            const float du = bu - u0, dv = bv - v0, dob = bo - o0;
            const long iv0 = to_isize(v0), iu0 = to_isize(u0);
            const size_t io0 = (size_t)to_isize(o0);
            for (int a = 0; a < 2; ++a) {
                const long iv = iv0 + a;
                const float wv = a ? dv : 1.0f - dv;
                if (iv < 0 || iv >= NC) continue;
                for (int b = 0; b < 2; ++b) {
                    const long iu = iu0 + b;
                    const float wu = b ? du : 1.0f - du;
                    if (iu < 0 || iu >= NC) continue;
                    for (int o = 0; o < 2; ++o) {
                        const size_t io = (io0 + (size_t)o) % NO;
                        const float wo = o ? dob : 1.0f - dob;
                        const size_t bin = ((size_t)iv * NC + (size_t)iu) * NO + io;
                        hist[bin] = hist[bin] + m * wv * wu * wo;
                    }
                }
            }
        }
    }
    for (int pass = 0; pass < 2; ++pass) {
        float s = -0.0f;
        for (int i = 0; i < ORC_DESC_LEN; ++i) s = s + hist[i] * hist[i];
        const float n = sqrtf(s);
        if (n > 0.0f)
            for (int i = 0; i < ORC_DESC_LEN; ++i) hist[i] = hist[i] / n;
        if (pass == 0)
            for (int i = 0; i < ORC_DESC_LEN; ++i) hist[i] = fminf(hist[i], 0.2f);
    }
}

/* Gradient table of one layer (GradTable::new, features.rs:340-352): border
 * entries stay (0, 0) and are never read. */
static void grad_table(const float* img, size_t w, size_t h, float* mag, float* ang) {
    for (size_t i = 0; i < w * h; ++i) {
        mag[i] = 0.0f;
        ang[i] = 0.0f;
    }
    for (size_t y = 1; y + 1 < h; ++y)
        for (size_t x = 1; x + 1 < w; ++x) pixel_grad(img, w, x, y, &mag[y * w + x], &ang[y * w + x]);
}

size_t orc_dominant_orientations(const float* img, size_t w, size_t h, float x, float y,
                                 float sigma, float* out) {
    float* mag = (float*)malloc(sizeof(float) * w * h);
    float* ang = (float*)malloc(sizeof(float) * w * h);
    if (!mag || !ang) {
        free(mag);
        free(ang);
        return 0;
    }
    grad_table(img, w, h, mag, ang);
    const size_t n = orientations_with(w, h, mag, ang, x, y, sigma, out);
    free(mag);
    free(ang);
    return n;
}

void orc_describe(const float* img, size_t w, size_t h, float x, float y, float sigma,
                  float angle, float* desc) {
    float* mag = (float*)malloc(sizeof(float) * w * h);
    float* ang = (float*)malloc(sizeof(float) * w * h);
    if (!mag || !ang) {
        free(mag);
        free(ang);
        return;
    }
    grad_table(img, w, h, mag, ang);
    describe_with(w, h, mag, ang, x, y, sigma, angle, desc);
    free(mag);
    free(ang);
}

/* refine_extremum (features.rs:430-485). The 3x3 inverse is the cofactor
 * form (the reference calls nalgebra's try_inverse; its exact bits are not
 * available here, so this restatement fixes its own). */
typedef struct {
    size_t xi, yi, layer;
    float x, y, s, value;
} refined_t;

static int inverse3(const float* m, float* inv) {
    const float m11 = m[0], m12 = m[1], m13 = m[2];
    const float m21 = m[3], m22 = m[4], m23 = m[5];
    const float m31 = m[6], m32 = m[7], m33 = m[8];
    const float c11 = m22 * m33 - m32 * m23;
    const float c21 = m21 * m33 - m31 * m23;
    const float c31 = m21 * m32 - m31 * m22;
    const float det = m11 * c11 - m12 * c21 + m13 * c31;
    if (det == 0.0f) return 0;
    inv[0] = c11 / det;
    inv[1] = (m13 * m32 - m33 * m12) / det;
    inv[2] = (m12 * m23 - m22 * m13) / det;
    inv[3] = -c21 / det;
    inv[4] = (m11 * m33 - m31 * m13) / det;
    inv[5] = (m13 * m21 - m23 * m11) / det;
    inv[6] = c31 / det;
    inv[7] = (m12 * m31 - m32 * m11) / det;
    inv[8] = (m11 * m22 - m21 * m12) / det;
    return 1;
}

static int refine_extremum(float* const* dog, size_t w, size_t h, size_t x, size_t y, size_t l,
                           size_t s, refined_t* out) {
    for (int it = 0; it < 5; ++it) {
        const float* dm = dog[l - 1];
        const float* d0 = dog[l];
        const float* dp = dog[l + 1];
#define AT(d, xx, yy) ((d)[(yy) * w + (xx)])
        const float v = AT(d0, x, y);
        const float g[3] = {0.5f * (AT(d0, x + 1, y) - AT(d0, x - 1, y)),
                            0.5f * (AT(d0, x, y + 1) - AT(d0, x, y - 1)),
                            0.5f * (AT(dp, x, y) - AT(dm, x, y))};
        const float hxx = AT(d0, x + 1, y) + AT(d0, x - 1, y) - 2.0f * v;
        const float hyy = AT(d0, x, y + 1) + AT(d0, x, y - 1) - 2.0f * v;
        const float hss = AT(dp, x, y) + AT(dm, x, y) - 2.0f * v;
        const float hxy = 0.25f * (AT(d0, x + 1, y + 1) - AT(d0, x - 1, y + 1) -
                                   AT(d0, x + 1, y - 1) + AT(d0, x - 1, y - 1));
        const float hxs =
            0.25f * (AT(dp, x + 1, y) - AT(dp, x - 1, y) - AT(dm, x + 1, y) + AT(dm, x - 1, y));
        const float hys =
            0.25f * (AT(dp, x, y + 1) - AT(dp, x, y - 1) - AT(dm, x, y + 1) + AT(dm, x, y - 1));
#undef AT
        const float hm[9] = {hxx, hxy, hxs, hxy, hyy, hys, hxs, hys, hss};
        float inv[9];
        if (!inverse3(hm, inv)) return 0;
        float delta[3];
        for (int i = 0; i < 3; ++i) {
            const float p = inv[3 * i] * g[0] + inv[3 * i + 1] * g[1] + inv[3 * i + 2] * g[2];
            delta[i] = -p;
        }
        if (fabsf(delta[0]) <= 0.5f && fabsf(delta[1]) <= 0.5f && fabsf(delta[2]) <= 0.5f) {
            out->xi = x;
            out->yi = y;
            out->layer = l;
            out->x = (float)x + delta[0];
            out->y = (float)y + delta[1];
            out->s = (float)l + delta[2];
            out->value = v + 0.5f * (g[0] * delta[0] + g[1] * delta[1] + g[2] * delta[2]);
            return 1;
        }
        if (!isfinite(delta[0]) || !isfinite(delta[1]) || !isfinite(delta[2])) return 0;
        const long nx = (long)x + to_isize(roundf(delta[0]));
        const long ny = (long)y + to_isize(roundf(delta[1]));
        const long nl = (long)l + to_isize(roundf(delta[2]));
        if (nx < 1 || ny < 1 || nl < 1 || nx >= (long)w - 1 || ny >= (long)h - 1 ||
            nl > (long)s)
            return 0;
        x = (size_t)nx;
        y = (size_t)ny;
        l = (size_t)nl;
    }
    return 0;
}

/* ---- detect_core / detect_and_describe ------------------------------------ */

typedef struct {
    orc_keypoint kp;
    float desc[ORC_DESC_LEN];
    uint64_t order; /* emission order, for the stable sort */
} feat_t;

typedef struct {
    feat_t* v;
    size_t n, cap;
} feat_vec;

static int fv_push(feat_vec* fv, const feat_t* f) {
    if (fv->n == fv->cap) {
        size_t nc = fv->cap ? 2 * fv->cap : 256;
        feat_t* nv = (feat_t*)realloc(fv->v, nc * sizeof(feat_t));
        if (!nv) return -1;
        fv->v = nv;
        fv->cap = nc;
    }
    fv->v[fv->n++] = *f;
    return 0;
}

/* |response| descending by total order (abs clears the sign, so the bit
 * pattern orders it), ties by emission order: the reference's stable
 * sort_by on total_cmp (features.rs:778). */
static int cmp_feat(const void* pa, const void* pb) {
    const feat_t* a = (const feat_t*)pa;
    const feat_t* b = (const feat_t*)pb;
    uint32_t ka, kb;
    const float ra = fabsf(a->kp.response), rb = fabsf(b->kp.response);
    memcpy(&ka, &ra, 4);
    memcpy(&kb, &rb, 4);
    if (ka != kb) return ka > kb ? -1 : 1;
    return a->order < b->order ? -1 : (a->order > b->order);
}

/* detect_core (features.rs:650-781). Returns 0 and the features in fv. */
static int detect_core(const float* img, size_t w0, size_t h0, const orc_detector_cfg* cfg,
                       float in_sigma, feat_vec* fv) {
    const size_t s = cfg->scales;
    const float kstep = powf(2.0f, 1.0f / (float)s);
    float sig0 = cfg->sigma0 * cfg->sigma0 - in_sigma * in_sigma;
    sig0 = sig0 > 0.01f ? sig0 : 0.01f;
    size_t w = w0, h = h0;
    float* base = (float*)malloc(sizeof(float) * (w * h > 0 ? w * h : 1));
    if (!base || orc_gaussian_blur(img, w, h, sqrtf(sig0), base)) {
        free(base);
        return -1;
    }
    const size_t ng = s + 3, nd = s + 2;
    float** gauss = (float**)calloc(ng, sizeof(float*));
    float** dog = (float**)calloc(nd, sizeof(float*));
    float** gmag = (float**)calloc(ng, sizeof(float*));
    float** gang = (float**)calloc(ng, sizeof(float*));
    int rc = 0;
    uint64_t order = 0;
    float* angles = (float*)malloc(sizeof(float) * 36);
    const float edge_thr = powi_f(cfg->edge_ratio + 1.0f, 2) / cfg->edge_ratio;
    for (size_t o = 0; o < cfg->octaves && rc == 0; ++o) {
        if (w < 16 || h < 16) break;
        const size_t np = w * h;
        for (size_t i = 0; i < ng; ++i) gauss[i] = (float*)malloc(sizeof(float) * np);
        for (size_t i = 0; i < nd; ++i) dog[i] = (float*)malloc(sizeof(float) * np);
        for (size_t i = 0; i < ng; ++i) {
            gmag[i] = (float*)malloc(sizeof(float) * np);
            gang[i] = (float*)malloc(sizeof(float) * np);
        }
        memcpy(gauss[0], base, sizeof(float) * np);
        for (size_t i = 1; i < ng; ++i) {
            const float prev = cfg->sigma0 * powi_f(kstep, (int)i - 1);
            const float cur = prev * kstep;
            if (orc_gaussian_blur(gauss[i - 1], w, h, sqrtf(cur * cur - prev * prev), gauss[i]))
                rc = -1;
        }
        for (size_t i = 0; i < nd; ++i)
            for (size_t p = 0; p < np; ++p) dog[i][p] = gauss[i + 1][p] - gauss[i][p];
        for (size_t i = 1; i <= s; ++i) grad_table(gauss[i], w, h, gmag[i], gang[i]);
        const float scale = (float)((size_t)1 << o);
        for (size_t l = 1; l <= s && rc == 0; ++l) {
            const float *d0 = dog[l - 1], *d1 = dog[l], *d2 = dog[l + 1];
            for (size_t y = 1; y + 1 < h; ++y) {
                for (size_t x = 1; x + 1 < w; ++x) {
                    const float v = d1[y * w + x];
                    if (fabsf(v) < cfg->contrast) continue;
                    int is_max = 1, is_min = 1;
                    const float* ds[3] = {d0, d1, d2};
                    for (int k = 0; k < 3 && (is_max || is_min); ++k)
                        for (size_t yy = y - 1; yy <= y + 1; ++yy)
                            for (size_t xx = x - 1; xx <= x + 1; ++xx) {
                                if (k == 1 && xx == x && yy == y) continue;
                                const float u = ds[k][yy * w + xx];
                                is_max &= v > u;
                                is_min &= v < u;
                            }
                    if (!(is_max || is_min)) continue;
                    refined_t r;
                    if (!refine_extremum(dog, w, h, x, y, l, s, &r)) continue;
                    if (fabsf(r.value) < cfg->contrast) continue;
                    const float* d = dog[r.layer];
                    const size_t xi = r.xi, yi = r.yi;
                    const float c = d[yi * w + xi];
                    const float dxx = d[yi * w + xi + 1] + d[yi * w + xi - 1] - 2.0f * c;
                    const float dyy = d[(yi + 1) * w + xi] + d[(yi - 1) * w + xi] - 2.0f * c;
                    const float dxy =
                        0.25f * (d[(yi + 1) * w + xi + 1] - d[(yi + 1) * w + xi - 1] -
                                 d[(yi - 1) * w + xi + 1] + d[(yi - 1) * w + xi - 1]);
                    const float tr = dxx + dyy;
                    const float det = dxx * dyy - dxy * dxy;
                    if (det <= 0.0f || tr * tr / det >= edge_thr) continue;
                    const float sig_oct = cfg->sigma0 * orc_powf(kstep, r.s);
                    const size_t na = orientations_with(w, h, gmag[r.layer], gang[r.layer], r.x,
                                                        r.y, sig_oct, angles);
                    for (size_t a = 0; a < na; ++a) {
                        feat_t f;
                        f.kp.x = r.x * scale;
                        f.kp.y = r.y * scale;
                        f.kp.sigma = sig_oct * scale;
                        f.kp.response = r.value;
                        f.kp.angle = angles[a];
                        describe_with(w, h, gmag[r.layer], gang[r.layer], r.x, r.y, sig_oct,
                                      angles[a], f.desc);
                        f.order = order++;
                        if (fv_push(fv, &f)) rc = -1;
                    }
                }
            }
        }
        /* base = gauss[s].downsample() */
        const size_t W = w / 2, H = h / 2;
        float* nb = (float*)malloc(sizeof(float) * (W * H > 0 ? W * H : 1));
        orc_downsample(gauss[s], w, h, nb);
        free(base);
        base = nb;
        w = W;
        h = H;
        for (size_t i = 0; i < ng; ++i) {
            free(gauss[i]);
            free(gmag[i]);
            free(gang[i]);
            gauss[i] = gmag[i] = gang[i] = NULL;
        }
        for (size_t i = 0; i < nd; ++i) {
            free(dog[i]);
            dog[i] = NULL;
        }
    }
    free(angles);
    free(base);
    free(gauss);
    free(dog);
    free(gmag);
    free(gang);
    if (rc) return rc;
    qsort(fv->v, fv->n, sizeof(feat_t), cmp_feat);
    if (fv->n > cfg->max_features) fv->n = cfg->max_features;
    return 0;
}

void orc_detector_cfg_default(orc_detector_cfg* c) {
    c->octaves = 4;
    c->scales = 3;
    c->sigma0 = 1.6f;
    c->contrast = 0.01f;
    c->edge_ratio = 10.0f;
    c->max_features = 8192;
    c->upscale_below = 800;
    c->upscale_fill = 0;
}

/* detect_and_describe (features.rs:594-648). Writes up to cap features
 * (keypoint + 128 f32 each) and returns the total count, or -1 on an
 * allocation failure. */
long orc_detect_and_describe(const float* img, size_t w, size_t h, const orc_detector_cfg* cfg,
                             orc_keypoint* kps, float* desc, size_t cap) {
    feat_vec feats = {0}, extra = {0};
    if (cfg->scales == 0 || detect_core(img, w, h, cfg, 0.5f, &feats)) {
        free(feats.v);
        return cfg->scales == 0 ? 0 : -1;
    }
    const size_t longer = w > h ? w : h;
    size_t total = feats.n;
    feat_t** added = NULL;
    size_t n_added = 0;
    if (!(cfg->upscale_below == 0 || longer >= cfg->upscale_below ||
          feats.n >= cfg->max_features || (!cfg->upscale_fill && feats.n * 2 >= cfg->max_features))) {
        float* up = (float*)malloc(sizeof(float) * 4 * w * h);
        if (!up) {
            free(feats.v);
            return -1;
        }
        orc_upsample2(img, w, h, up);
        const int rc = detect_core(up, 2 * w, 2 * h, cfg, 1.0f, &extra);
        free(up);
        if (rc) {
            free(feats.v);
            free(extra.v);
            return -1;
        }
        const size_t need = cfg->max_features - feats.n;
        added = (feat_t**)malloc(sizeof(feat_t*) * (extra.n ? extra.n : 1));
        /* The reference looks up the original features in a 4-pixel grid around
         * the candidate (features.rs:612-639); any original within distance 1
         * lies in the 3 x 3 cells it visits, so scanning all of them is the
         * same test. */
        for (size_t e = 0; e < extra.n && n_added < need; ++e) {
            feat_t* f = &extra.v[e];
            f->kp.x = f->kp.x / 2.0f - 0.25f;
            f->kp.y = f->kp.y / 2.0f - 0.25f;
            f->kp.sigma = f->kp.sigma / 2.0f;
            int dup = 0;
            for (size_t k = 0; k < feats.n && !dup; ++k) {
                const float s = feats.v[k].kp.sigma;
                const float ra = s / f->kp.sigma, rb = f->kp.sigma / s;
                const float r = fmaxf(ra, rb);
                dup = hypot_f(feats.v[k].kp.x - f->kp.x, feats.v[k].kp.y - f->kp.y) <= 1.0f &&
                      r <= 1.5f;
            }
            if (!dup) added[n_added++] = f;
        }
        total += n_added;
    }
    for (size_t k = 0; k < total && k < cap; ++k) {
        const feat_t* f = k < feats.n ? &feats.v[k] : added[k - feats.n];
        if (kps) kps[k] = f->kp;
        if (desc) memcpy(desc + k * ORC_DESC_LEN, f->desc, sizeof(float) * ORC_DESC_LEN);
    }
    free(added);
    free(feats.v);
    free(extra.v);
    return (long)total;
}
