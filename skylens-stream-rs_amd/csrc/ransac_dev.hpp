// Device-side FP64 geometry of ransac_fundamental (gfx950, one lane per call).
//
// Reference: crates/core/src/matching.rs:217-1104 and
// crates/core/src/two_view.rs:605-642. Every routine performs the same
// floating-point operations in the same order as the host restatement, with
// no FMA contraction, so a lane's F, counts and masks are bit-identical to
// it. The eigen/SVD solvers are the project's own (cyclic Jacobi for the
// symmetric 9x9 problem, one-sided Jacobi for the 3x3 rank-2 projection):
// nalgebra's routines (the reference's) are not part of this build.
//
// Point sets are read through a selector: all points, a bitset over the
// n correspondences, or eight gathered samples, always in index order.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

namespace skh {
namespace rdev {

// ------------------------------------------------------ compile-time loops

template <int Base, class F, int... I>
__device__ __forceinline__ void sfor_impl(F& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, Base + I>{}), ...);
}

// f(std::integral_constant<int, i>) for i = B .. E-1, expanded at compile
// time so that every array index in f is a constant (register-resident).
template <int B, int E, class F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (E > B) sfor_impl<B>(f, std::make_integer_sequence<int, E - B>{});
}

// ------------------------------------------------------------ point sources

struct PtsAll {
    const double2* x1;
    const double2* x2;
    uint32_t n;
    __device__ bool take(uint32_t) const { return true; }
    __device__ double2 p(uint32_t i) const { return x1[i]; }
    __device__ double2 q(uint32_t i) const { return x2[i]; }
};

struct PtsMask {
    const double2* x1;
    const double2* x2;
    uint32_t n;
    const uint32_t* bits;
    __device__ bool take(uint32_t i) const { return (bits[i >> 5] >> (i & 31)) & 1u; }
    __device__ double2 p(uint32_t i) const { return x1[i]; }
    __device__ double2 q(uint32_t i) const { return x2[i]; }
};

struct Pts8 {
    double2 a[8];
    double2 b[8];
    static constexpr uint32_t n = 8;
    __device__ bool take(uint32_t) const { return true; }
    __device__ double2 p(uint32_t i) const { return a[i]; }
    __device__ double2 q(uint32_t i) const { return b[i]; }
};

// Eight samples read through their indices (the indices live in LDS).
struct PtsIdx8 {
    const double2* x1;
    const double2* x2;
    const uint32_t* idx;
    static constexpr uint32_t n = 8;
    __device__ bool take(uint32_t) const { return true; }
    __device__ double2 p(uint32_t i) const { return x1[idx[i]]; }
    __device__ double2 q(uint32_t i) const { return x2[idx[i]]; }
};

// ------------------------------------------------------------------ 3x3 ops

struct M3 {
    double m[9];
};

__device__ __forceinline__ void mv3(const double* m, const double* x, double* y) {
    double a = m[0] * x[0] + m[1] * x[1] + m[2] * x[2];
    double b = m[3] * x[0] + m[4] * x[1] + m[5] * x[2];
    double c = m[6] * x[0] + m[7] * x[1] + m[8] * x[2];
    y[0] = a;
    y[1] = b;
    y[2] = c;
}

__device__ __forceinline__ void mtv3(const double* m, const double* x, double* y) {
    double a = m[0] * x[0] + m[3] * x[1] + m[6] * x[2];
    double b = m[1] * x[0] + m[4] * x[1] + m[7] * x[2];
    double c = m[2] * x[0] + m[5] * x[1] + m[8] * x[2];
    y[0] = a;
    y[1] = b;
    y[2] = c;
}

__device__ __forceinline__ void mm3(const double* A, const double* B, double* C) {
    double t[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            t[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
#pragma unroll
    for (int i = 0; i < 9; ++i) C[i] = t[i];
}

__device__ __forceinline__ void tr3(const double* A, double* T) {
    double t[9] = {A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8]};
#pragma unroll
    for (int i = 0; i < 9; ++i) T[i] = t[i];
}

__device__ __forceinline__ double fro3(const double* m) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double col = m[c] * m[c] + m[3 + c] * m[3 + c] + m[6 + c] * m[6 + c];
        s = s + col;
    }
    return sqrt(s);
}

__device__ __forceinline__ double norm2(double x, double y) { return sqrt(x * x + y * y); }

__device__ __forceinline__ bool inv3(const double* m, double* out) {
    double m11 = m[4] * m[8] - m[5] * m[7];
    double m12 = m[3] * m[8] - m[5] * m[6];
    double m13 = m[3] * m[7] - m[4] * m[6];
    double det = m[0] * m11 - m[1] * m12 + m[2] * m13;
    if (det == 0.0) return false;
    double o[9];
    o[0] = m11 / det;
    o[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    o[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    o[3] = -m12 / det;
    o[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    o[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    o[6] = m13 / det;
    o[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    o[8] = (m[0] * m[4] - m[1] * m[3]) / det;
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = o[i];
    return true;
}

// ---------------------------------------------------------- Jacobi eigen 9x9

// Symmetric matrix stored as its upper triangle, index of (r, c), r <= c.
__host__ __device__ constexpr int ut(int r, int c) {
    return r <= c ? r * 9 - r * (r - 1) / 2 + (c - r) : c * 9 - c * (c - 1) / 2 + (r - c);
}

// Rotation (c, s, t) annihilating a[p][q]; false when none is due (a zero
// element, or after the fourth sweep one negligible against both diagonal
// entries, which the sweep then just sets to zero).
__device__ __forceinline__ bool jacobi_params(double app, double aqq, double apq, int sweep,
                                              double& c, double& s, double& t) {
    c = 1.0;
    s = 0.0;
    t = 0.0;
    if (apq == 0.0) return false;
    double g = 100.0 * fabs(apq);
    if (sweep > 3 && fabs(app) + g == fabs(app) && fabs(aqq) + g == fabs(aqq)) return false;
    double h = aqq - app;
    double tt;
    if (fabs(h) + g == fabs(h)) {
        tt = apq / h;
    } else {
        double theta = 0.5 * h / apq;
        tt = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
        if (theta < 0.0) tt = -tt;
    }
    double cc = 1.0 / sqrt(1.0 + tt * tt);
    c = cc;
    s = tt * cc;
    t = tt;
    return true;
}

// Rotates the upper triangle `a` (45) by (c, s, t) in the (p, q) plane.
template <int p, int q>
__device__ __forceinline__ void jacobi_rotate_a(double* a, double c, double s, double t) {
    const double apq = a[ut(p, q)];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        if (k == p || k == q) continue;
        double akp = a[ut(k, p)], akq = a[ut(k, q)];
        a[ut(k, p)] = c * akp - s * akq;
        a[ut(k, q)] = s * akp + c * akq;
    }
    a[ut(p, p)] = a[ut(p, p)] - t * apq;
    a[ut(q, q)] = a[ut(q, q)] + t * apq;
}

template <int R>
__device__ __forceinline__ void jacobi_rotate_v(double (*v)[9], int p, int q, double c, double s) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
        double vkp = v[k][p], vkq = v[k][q];
        v[k][p] = c * vkp - s * vkq;
        v[k][q] = s * vkp + c * vkq;
    }
}

// Cyclic Jacobi on the upper triangle `a` (45), accumulating R rows of the
// eigenvector matrix starting at row r0 (v[i] = row r0 + i; rows past 8 are
// padding). Same rotation order and arithmetic as the host's cyclic Jacobi;
// rows of the eigenvector matrix evolve independently, so a pair of lanes
// can split them (R = 5) while both rotate `a`, which keeps a and v in
// registers.
template <int R>
__device__ __forceinline__ void sym_eigen9_rows(double* a, int r0, double (*v)[9]) {
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j) v[i][j] = (r0 + i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0;
#pragma unroll
        for (int p = 0; p < 8; ++p)
#pragma unroll
            for (int q = p + 1; q < 9; ++q) off = off + fabs(a[ut(p, q)]);
        if (off == 0.0) break;
#define SKH_ROT(P, Q)                                                                 \
    {                                                                                 \
        double c, s, t;                                                               \
        if (jacobi_params(a[ut(P, P)], a[ut(Q, Q)], a[ut(P, Q)], sweep, c, s, t)) {   \
            jacobi_rotate_a<P, Q>(a, c, s, t);                                        \
            jacobi_rotate_v<R>(v, P, Q, c, s);                                        \
        }                                                                             \
        a[ut(P, Q)] = 0.0;                                                            \
    }
        SKH_ROT(0, 1) SKH_ROT(0, 2) SKH_ROT(0, 3) SKH_ROT(0, 4) SKH_ROT(0, 5) SKH_ROT(0, 6)
        SKH_ROT(0, 7) SKH_ROT(0, 8) SKH_ROT(1, 2) SKH_ROT(1, 3) SKH_ROT(1, 4) SKH_ROT(1, 5)
        SKH_ROT(1, 6) SKH_ROT(1, 7) SKH_ROT(1, 8) SKH_ROT(2, 3) SKH_ROT(2, 4) SKH_ROT(2, 5)
        SKH_ROT(2, 6) SKH_ROT(2, 7) SKH_ROT(2, 8) SKH_ROT(3, 4) SKH_ROT(3, 5) SKH_ROT(3, 6)
        SKH_ROT(3, 7) SKH_ROT(3, 8) SKH_ROT(4, 5) SKH_ROT(4, 6) SKH_ROT(4, 7) SKH_ROT(4, 8)
        SKH_ROT(5, 6) SKH_ROT(5, 7) SKH_ROT(5, 8) SKH_ROT(6, 7) SKH_ROT(6, 8) SKH_ROT(7, 8)
#undef SKH_ROT
    }
}

// One lane, all nine rows (v: 9x9 row-major).
__device__ __forceinline__ void sym_eigen9(double* a, double* v) {
    sym_eigen9_rows<9>(a, 0, reinterpret_cast<double (*)[9]>(v));
}

// --------------------------------------------------------------- 3x3 rank 2

__device__ __forceinline__ void rank2_recompose(const double* F, double* out) {
    double w[9], v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
    for (int i = 0; i < 9; ++i) w[i] = F[i];
    for (int sweep = 0; sweep < 30; ++sweep) {
        bool rotated = false;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int p = r == 2 ? 1 : 0, q = r == 0 ? 1 : 2;
            double alpha = w[p] * w[p] + w[3 + p] * w[3 + p] + w[6 + p] * w[6 + p];
            double beta = w[q] * w[q] + w[3 + q] * w[3 + q] + w[6 + q] * w[6 + q];
            double gamma = w[p] * w[q] + w[3 + p] * w[3 + q] + w[6 + p] * w[6 + q];
            double c = 1.0, s = 0.0;
            if (gamma != 0.0 && fabs(gamma) > 1e-15 * sqrt(alpha * beta)) {
                double zeta = (beta - alpha) / (2.0 * gamma);
                double t = 1.0 / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                if (zeta < 0.0) t = -t;
                c = 1.0 / sqrt(1.0 + t * t);
                s = c * t;
                rotated = true;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double wp = w[3 * k + p], wq = w[3 * k + q];
                w[3 * k + p] = c * wp - s * wq;
                w[3 * k + q] = s * wp + c * wq;
                double vp = v[3 * k + p], vq = v[3 * k + q];
                v[3 * k + p] = c * vp - s * vq;
                v[3 * k + q] = s * vp + c * vq;
            }
        }
        if (!rotated) break;
    }
    double sig[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) sig[c] = w[c] * w[c] + w[3 + c] * w[3 + c] + w[6 + c] * w[6 + c];
    int m = 0;
    if (sig[1] < sig[m]) m = 1;
    if (sig[2] < sig[m]) m = 2;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                if (i == m) continue;
                acc = acc + w[3 * r + i] * v[3 * c + i];
            }
            out[3 * r + c] = acc;
        }
}

__device__ __forceinline__ bool rank2_unit(const double* f, double* out) {
    double g[9];
    rank2_recompose(f, g);
    double n = fro3(g);
    if (!(isfinite(n) && n > 0.0)) return false;
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = g[i] / n;
    return true;
}

// ------------------------------------------------------------- geometry ops

template <class S>
__device__ void normalizer(const S& s, bool second, double* T, uint32_t len) {
    double n = (double)len;
    double cx = 0.0, cy = 0.0;
    for (uint32_t i = 0; i < s.n; ++i) {
        if (!s.take(i)) continue;
        double2 v = second ? s.q(i) : s.p(i);
        cx = cx + v.x;
        cy = cy + v.y;
    }
    cx = cx / n;
    cy = cy / n;
    double d = 0.0;
    for (uint32_t i = 0; i < s.n; ++i) {
        if (!s.take(i)) continue;
        double2 v = second ? s.q(i) : s.p(i);
        d = d + norm2(v.x - cx, v.y - cy);
    }
    d = d / n;
    double sc = d > 0.0 ? 1.4142135623730951 / d : 1.0;
    T[0] = sc; T[1] = 0.0; T[2] = -sc * cx;
    T[3] = 0.0; T[4] = sc; T[5] = -sc * cy;
    T[6] = 0.0; T[7] = 0.0; T[8] = 1.0;
}

template <class S>
__device__ bool nearly_collinear(const S& s, bool second, uint32_t len) {
    double n = (double)(len > 0 ? len : 1);
    double cx = 0.0, cy = 0.0;
    for (uint32_t i = 0; i < s.n; ++i) {
        if (!s.take(i)) continue;
        double2 v = second ? s.q(i) : s.p(i);
        cx = cx + v.x;
        cy = cy + v.y;
    }
    cx = cx / n;
    cy = cy / n;
    double sxx = 0.0, sxy = 0.0, syy = 0.0;
    for (uint32_t i = 0; i < s.n; ++i) {
        if (!s.take(i)) continue;
        double2 v = second ? s.q(i) : s.p(i);
        double dx = v.x - cx, dy = v.y - cy;
        sxx += dx * dx;
        sxy += dx * dy;
        syy += dy * dy;
    }
    sxx = sxx / n;
    sxy = sxy / n;
    syy = syy / n;
    double tr = sxx + syy;
    double dd = sxx - syy;
    double disc = sqrt(dd * dd + 4.0 * sxy * sxy);
    double hi = 0.5 * (tr + disc);
    double lo = 0.5 * (tr - disc);
    if (!(lo >= 0.0)) lo = 0.0;
    double slo = sqrt(lo), shi = sqrt(hi);
    return !(shi > 0.0 && slo > 0.02 * shi);
}

template <class S>
__device__ uint32_t count_taken(const S& s) {
    uint32_t k = 0;
    for (uint32_t i = 0; i < s.n; ++i) k += s.take(i) ? 1u : 0u;
    return k;
}

template <class S>
__device__ bool all_finite(const S& s) {
    for (uint32_t i = 0; i < s.n; ++i) {
        if (!s.take(i)) continue;
        double2 a = s.p(i), b = s.q(i);
        if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(b.x) || !isfinite(b.y)) return false;
    }
    return true;
}

// From the centroid sums (c: x1.x, x1.y, x2.x, x2.y) and the second-pass
// sums (q: s1xx, s1xy, s1yy, s2xx, s2xy, s2yy, d1, d2) of len points: the
// Hartley normalisers t1, t2 (matching.rs:498-505) and whether the points of
// either image are nearly collinear (matching.rs:224-247).
__device__ __forceinline__ bool collinear_finish(double c1x, double c1y, double c2x, double c2y,
                                                 const double* q, uint32_t len, double* t1,
                                                 double* t2) {
    const double nn = (double)len, nc = (double)(len > 0 ? len : 1);
    bool collinear = false;
#pragma unroll
    for (int im = 0; im < 2; ++im) {
        double sxx = q[3 * im] / nc, sxy = q[3 * im + 1] / nc, syy = q[3 * im + 2] / nc;
        double tr = sxx + syy;
        double dd = sxx - syy;
        double disc = sqrt(dd * dd + 4.0 * sxy * sxy);
        double hi = 0.5 * (tr + disc);
        double lo = 0.5 * (tr - disc);
        if (!(lo >= 0.0)) lo = 0.0;
        double slo = sqrt(lo), shi = sqrt(hi);
        if (!(shi > 0.0 && slo > 0.02 * shi)) collinear = true;
    }
    const double m1 = q[6] / nn, m2 = q[7] / nn;
    const double sc1 = m1 > 0.0 ? 1.4142135623730951 / m1 : 1.0;
    const double sc2 = m2 > 0.0 ? 1.4142135623730951 / m2 : 1.0;
    t1[0] = sc1; t1[1] = 0.0; t1[2] = -sc1 * c1x;
    t1[3] = 0.0; t1[4] = sc1; t1[5] = -sc1 * c1y;
    t1[6] = 0.0; t1[7] = 0.0; t1[8] = 1.0;
    t2[0] = sc2; t2[1] = 0.0; t2[2] = -sc2 * c2x;
    t2[3] = 0.0; t2[4] = sc2; t2[5] = -sc2 * c2y;
    t2[6] = 0.0; t2[7] = 0.0; t2[8] = 1.0;
    return collinear;
}

// Collinearity tests and Hartley normalisers of both images in two passes.
// The host computes nearly_collinear(x1), nearly_collinear(x2),
// normalizer(x1), normalizer(x2) one after another; each of their sums sees
// the same operands in the same order here, so every value is identical
// (axis_spread divides the centroid by max(len, 1), the normaliser by len:
// equal for len >= 1).
template <class S>
__device__ bool collinear_and_normalizers(const S& s, uint32_t len, double* t1, double* t2) {
    double c1x = 0.0, c1y = 0.0, c2x = 0.0, c2y = 0.0;
    for (uint32_t i = 0; i < s.n; ++i) {
        if (!s.take(i)) continue;
        const double2 a = s.p(i), b = s.q(i);
        c1x = c1x + a.x;
        c1y = c1y + a.y;
        c2x = c2x + b.x;
        c2y = c2y + b.y;
    }
    const double nc = (double)(len > 0 ? len : 1);
    c1x = c1x / nc;
    c1y = c1y / nc;
    c2x = c2x / nc;
    c2y = c2y / nc;
    double q[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = 0; i < s.n; ++i) {
        if (!s.take(i)) continue;
        const double2 a = s.p(i), b = s.q(i);
        const double ax = a.x - c1x, ay = a.y - c1y, bx = b.x - c2x, by = b.y - c2y;
        q[0] += ax * ax;
        q[1] += ax * ay;
        q[2] += ay * ay;
        q[3] += bx * bx;
        q[4] += bx * by;
        q[5] += by * by;
        q[6] = q[6] + norm2(ax, ay);
        q[7] = q[7] + norm2(bx, by);
    }
    return collinear_finish(c1x, c1y, c2x, c2y, q, len, t1, t2);
}

// matching.rs:512-538 over the selected points (len of them): the finite
// and collinearity checks, the Hartley transforms and A^T A (upper triangle).
template <class S>
__device__ bool accumulate_8pt(const S& s, uint32_t len, double* t1, double* t2, double* ata) {
    if (len < 8 || !all_finite(s)) return false;
    if (collinear_and_normalizers(s, len, t1, t2)) return false;
#pragma unroll
    for (int i = 0; i < 45; ++i) ata[i] = 0.0;
    for (uint32_t i = 0; i < s.n; ++i) {
        if (!s.take(i)) continue;
        double2 p = s.p(i), q = s.q(i);
        double pa[3] = {p.x, p.y, 1.0}, pb[3] = {q.x, q.y, 1.0};
        double a[3], b[3];
        mv3(t1, pa, a);
        mv3(t2, pb, b);
        double row[9] = {b[0] * a[0], b[0] * a[1], b[0], b[1] * a[0], b[1] * a[1], b[1],
                         a[0], a[1], 1.0};
#pragma unroll
        for (int r = 0; r < 9; ++r)
#pragma unroll
            for (int c = r; c < 9; ++c) ata[ut(r, c)] = ata[ut(r, c)] + row[r] * row[c];
    }
    return true;
}

// Lanes 3g, 3g+1, 3g+2 form solver group g (g < kGroups); a group splits the
// nine eigenvector rows of a 9x9 eigen-solve three per lane while every lane
// of the group rotates the same matrix, which keeps both in registers.
constexpr int kGroups = 21;

__device__ __forceinline__ int solver_group() { return (int)threadIdx.x / 3; }
__device__ __forceinline__ int solver_part() { return (int)threadIdx.x - 3 * ((int)threadIdx.x / 3); }

// Eigen-solve of the upper triangle `a` (45, overwritten with the final
// rotated matrix) by a solver group; on return every lane of the group holds
// the eigenvector of the smallest diagonal entry (first minimum) in vk.
// All three lanes of the group must be active with the same `a`.
__device__ __forceinline__ int eigen_min_group(double* a, double* vk) {
    const int part = solver_part();
    const int base = 3 * solver_group();
    double v[3][9];
    sym_eigen9_rows<3>(a, 3 * part, v);
    // first minimum of the diagonal (the value is tracked so that no array
    // is indexed at run time)
    int k = 0;
    double dmin = a[ut(0, 0)];
#pragma unroll
    for (int i = 1; i < 9; ++i) {
        const double x = a[ut(i, i)];
        if (x < dmin) {
            dmin = x;
            k = i;
        }
    }
    double mine[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 9; ++j) acc = (j == k) ? v[i][j] : acc;
        mine[i] = acc;
    }
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int i = 0; i < 3; ++i) vk[3 * t + i] = __shfl(mine[i], base + t);
    return k;
}

// matching.rs:539-556 computed by a solver group (see eigen_min_group):
// the eigen-solve, the degeneracy test, the rank-2 projection and the
// denormalisation. ata is overwritten.
__device__ bool solve_8pt_group(double* ata, const double* t1, const double* t2, double* F) {
    double fn[9];
    eigen_min_group(ata, fn);
    // second smallest and largest of the nine eigenvalues
    double e1 = INFINITY, e0 = INFINITY, e8 = -INFINITY;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double x = ata[ut(i, i)];
        if (x < e0) {
            e1 = e0;
            e0 = x;
        } else if (x < e1) {
            e1 = x;
        }
        if (x > e8) e8 = x;
    }
    if (e1 <= 1e-10 * e8) return false;
    double r2[9];
    rank2_recompose(fn, r2);
    double t2t[9], tmp[9], f[9];
    tr3(t2, t2t);
    mm3(t2t, r2, tmp);
    mm3(tmp, t1, f);
    double nn = fro3(f);
    if (!(isfinite(nn) && nn > 0.0)) return false;
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = f[i] / nn;
    return true;
}

// matching.rs:561-573.
__device__ __forceinline__ double sampson_error(const double* F, double2 p, double2 q) {
    double a[3] = {p.x, p.y, 1.0}, b[3] = {q.x, q.y, 1.0};
    double fa[3], ftb[3];
    mv3(F, a, fa);
    mtv3(F, b, ftb);
    double e = b[0] * fa[0] + b[1] * fa[1] + b[2] * fa[2];
    double den = fa[0] * fa[0] + fa[1] * fa[1] + ftb[0] * ftb[0] + ftb[1] * ftb[1];
    if (den <= 0.0) return INFINITY;
    return e * e / den;
}

}  // namespace rdev
}  // namespace skh
