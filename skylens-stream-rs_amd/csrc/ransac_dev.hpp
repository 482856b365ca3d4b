// Device-side FP64 geometry of ransac_fundamental (gfx950, one lane per call).
//
// Reference: crates/core/src/matching.rs:217-1104 and
// crates/core/src/two_view.rs:605-642. Every routine performs the same
// floating-point operations in the same order as the host restatement, with
// no FMA contraction, so a lane's F, counts and masks are bit-identical to
// it. The eigen/SVD solvers are the project's own (Householder
// tridiagonalisation and implicit QL for the symmetric 9x9 problem,
// one-sided Jacobi for the 3x3 rank-2 projection): nalgebra's routines (the
// reference's) are not part of this build.
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

// ------------------------------------------------------- symmetric eigen 9x9
//
// Householder tridiagonalisation, then implicit QL with shifts, the same
// operations in the same order as the host restatement. The rows of the
// eigenvector matrix transform independently (every update multiplies it
// from the right), so a solver group of three lanes splits them three per
// lane while each lane reduces the same matrix: all state stays in
// registers and every array index is a compile-time constant. Scheduling
// barriers between the unrolled steps keep the scheduler from hoisting work
// across them, which would multiply the live registers.

// Symmetric matrix stored as its upper triangle, index of (r, c), r <= c.
__host__ __device__ constexpr int ut(int r, int c) {
    return r <= c ? r * 9 - r * (r - 1) / 2 + (c - r) : c * 9 - c * (c - 1) / 2 + (r - c);
}

// sqrt(a^2 + b^2), formed directly (the operands of the QL sweep stay far
// from overflow; an underflow to 0 is the sweep's r == 0 case).
__device__ __forceinline__ double hyp2(double a, double b) { return sqrt(a * a + b * b); }

// Tridiagonalises `a` (upper triangle, 45, overwritten); q: rows r0..r0+R-1
// of H_0 ... H_6; d: diagonal, e[i]: (i, i+1) coupling, e[8] = 0.
template <int R>
__device__ __forceinline__ void tridiag9_rows(double* a, int r0, double (*q)[9], double* d,
                                              double* e) {
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j) q[i][j] = (r0 + i == j) ? 1.0 : 0.0;
    sfor<0, 7>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        double xn2 = 0.0;
        sfor<k + 2, 9>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            xn2 = xn2 + a[ut(i, k)] * a[ut(i, k)];
        });
        const double x0 = a[ut(k + 1, k)];
        if (xn2 != 0.0) {
            const double nrm = sqrt(x0 * x0 + xn2);
            const double beta = x0 >= 0.0 ? -nrm : nrm;
            const double tau = (beta - x0) / beta;
            const double rs = 1.0 / (x0 - beta);
            double v[9], p[9], w[9];
            v[k + 1] = 1.0;
            sfor<k + 2, 9>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                v[i] = a[ut(i, k)] * rs;
            });
            sfor<k + 1, 9>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                double sum = 0.0;
                sfor<k + 1, 9>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    sum = sum + a[ut(i, j)] * v[j];
                });
                p[i] = tau * sum;
            });
            double vp = 0.0;
            sfor<k + 1, 9>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                vp = vp + v[i] * p[i];
            });
            const double kk = 0.5 * tau * vp;
            sfor<k + 1, 9>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                w[i] = p[i] - kk * v[i];
            });
            sfor<k + 1, 9>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                sfor<i, 9>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    a[ut(i, j)] = a[ut(i, j)] - (v[i] * w[j] + w[i] * v[j]);
                });
            });
            a[ut(k, k + 1)] = beta;
            sfor<k + 2, 9>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                a[ut(k, i)] = 0.0;
            });
#pragma unroll
            for (int r = 0; r < R; ++r) {
                double sum = 0.0;
                sfor<k + 1, 9>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    sum = sum + q[r][j] * v[j];
                });
                const double t = tau * sum;
                sfor<k + 1, 9>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    q[r][j] = q[r][j] - t * v[j];
                });
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    });
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] = a[ut(i, i)];
#pragma unroll
    for (int i = 0; i < 8; ++i) e[i] = a[ut(i, i + 1)];
    e[8] = 0.0;
}

// Eigenvalues of the symmetric tridiagonal (d, e) by implicit QL with
// shifts (tql1 scheme: tql2 without eigenvectors); each rotation's cosine
// and sine are g and f times 1 / r; at most 30 sweeps per eigenvalue.
__device__ __forceinline__ void tql9_values(double* d, double* e) {
    sfor<0, 9>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        for (int iter = 0; iter < 30; ++iter) {
            // first m >= l with a negligible e[m] (8 if none)
            int m = 8;
            sfor<l, 8>([&](auto mc) {
                constexpr int mm = decltype(mc)::value;
                const double dd = fabs(d[mm]) + fabs(d[mm + 1]);
                if (m == 8 && fabs(e[mm]) + dd == dd) m = mm;
            });
            if (m == l) break;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hyp2(g, 1.0);
            double dm = d[8];
#pragma unroll
            for (int j = l; j < 8; ++j) dm = (j == m) ? d[j] : dm;
            g = dm - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            bool underflow = false;
            sfor<0, 8 - l>([&](auto ic) {
                constexpr int i = 7 - decltype(ic)::value;
                if (i < m && !underflow) {
                    const double f = s * e[i], b = c * e[i];
                    r = hyp2(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] = d[i + 1] - p;
#pragma unroll
                        for (int j = l; j < 9; ++j) e[j] = (j == m) ? 0.0 : e[j];
                        underflow = true;
                    } else {
                        const double ri = 1.0 / r;
                        s = f * ri;
                        c = g * ri;
                        g = d[i + 1] - p;
                        r = (d[i] - g) * s + 2.0 * c * b;
                        p = s * r;
                        d[i + 1] = g + p;
                        g = c * r - b;
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            });
            if (underflow) continue;
            d[l] = d[l] - p;
            e[l] = g;
#pragma unroll
            for (int j = l; j < 9; ++j) e[j] = (j == m) ? 0.0 : e[j];
        }
    });
}

// x / max|x_i|, then divided by its 2-norm (sums in index order).
__device__ __forceinline__ void unit9(double* x) {
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) mx = fmax(mx, fabs(x[i]));
    const double sm = 1.0 / mx;
    double ss = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        x[i] = x[i] * sm;
        ss = ss + x[i] * x[i];
    }
    const double sn = 1.0 / sqrt(ss);
#pragma unroll
    for (int i = 0; i < 9; ++i) x[i] = x[i] * sn;
}

// Unit eigenvector of the symmetric tridiagonal (d, e) for its eigenvalue
// lam by inverse iteration (Wilkinson 1965, ch. 5.54; EISPACK tinvit):
// T - lam I = P L U by Gaussian elimination with partial pivoting (U with
// two superdiagonals), every pivot of U below eps3 = 2^-52 |T| in magnitude
// raised to eps3 with its sign (|T|: largest absolute row sum); then
// x1 = U^-1 (1, ..., 1), x2 = (P L U)^-1 unit(x1), x = unit(x2).
__device__ __forceinline__ void tinv9(const double* d, const double* e, double lam, double* x) {
    double tn = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double rs = fabs(d[i]);
        if (i < 8) rs = rs + fabs(e[i]);
        if (i > 0) rs = rs + fabs(e[i - 1]);
        tn = fmax(tn, rs);
    }
    const double eps3 = tn * 0x1p-52;
    double dd[9], du[8], du2[7], dl[8];
    bool swp[8];
#pragma unroll
    for (int i = 0; i < 9; ++i) dd[i] = d[i] - lam;
#pragma unroll
    for (int i = 0; i < 8; ++i) du[i] = dl[i] = e[i];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const bool keep = fabs(dd[i]) >= fabs(dl[i]);
        swp[i] = !keep;
        if (keep) {
            const double fact = dd[i] != 0.0 ? dl[i] / dd[i] : 0.0;
            dl[i] = fact;
            dd[i + 1] = dd[i + 1] - fact * du[i];
            if (i < 7) du2[i] = 0.0;
        } else {
            const double fact = dd[i] / dl[i];
            dd[i] = dl[i];
            dl[i] = fact;
            const double t = du[i];
            du[i] = dd[i + 1];
            dd[i + 1] = t - fact * dd[i + 1];
            if (i < 7) {
                du2[i] = du[i + 1];
                du[i + 1] = -fact * du[i + 1];
            }
        }
    }
    double ui[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double u = dd[i];
        if (fabs(u) < eps3) u = u >= 0.0 ? eps3 : -eps3;
        ui[i] = 1.0 / u;
    }
    double b[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) b[i] = 1.0;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        if (it == 1) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (swp[i]) {
                    const double t = b[i];
                    b[i] = b[i + 1];
                    b[i + 1] = t - dl[i] * b[i];
                } else {
                    b[i + 1] = b[i + 1] - dl[i] * b[i];
                }
            }
        }
        x[8] = b[8] * ui[8];
        x[7] = (b[7] - du[7] * x[8]) * ui[7];
#pragma unroll
        for (int i = 6; i >= 0; --i) x[i] = (b[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) * ui[i];
        unit9(x);
#pragma unroll
        for (int i = 0; i < 9; ++i) b[i] = x[i];
    }
}

// --------------------------------------------------------------- 3x3 rank 2

// One-sided Jacobi on W = F V (columns orthogonalised), the least-norm
// column dropped, F' = W' V^T. The rotations start from the identity (V0 =
// nullptr: W = F) or from the orthogonal basis V0 (row-major, column j =
// basis vector j: W = F V0, V = V0); Vout (if set) receives V.
__device__ __forceinline__ void rank2_recompose_v(const double* F, const double* V0, double* out,
                                                  double* Vout) {
    double w[9], v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (V0) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                w[3 * r + j] = F[3 * r] * V0[j] + F[3 * r + 1] * V0[3 + j] + F[3 * r + 2] * V0[6 + j];
#pragma unroll
        for (int i = 0; i < 9; ++i) v[i] = V0[i];
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) w[i] = F[i];
    }
    for (int sweep = 0; sweep < 30; ++sweep) {
        bool rotated = false;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int p = r == 2 ? 1 : 0, q = r == 0 ? 1 : 2;
            double alpha = w[p] * w[p] + w[3 + p] * w[3 + p] + w[6 + p] * w[6 + p];
            double beta = w[q] * w[q] + w[3 + q] * w[3 + q] + w[6 + q] * w[6 + q];
            double gamma = w[p] * w[q] + w[3 + p] * w[3 + q] + w[6 + p] * w[6 + q];
            double c = 1.0, s = 0.0;
            if (gamma != 0.0 && gamma * gamma > 1e-30 * (alpha * beta)) {
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
    if (Vout) {
#pragma unroll
        for (int i = 0; i < 9; ++i) Vout[i] = v[i];
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

__device__ __forceinline__ void rank2_recompose(const double* F, double* out) {
    rank2_recompose_v(F, nullptr, out, nullptr);
}

__device__ __forceinline__ bool rank2_unit_v(const double* f, const double* V0, double* out,
                                             double* Vout) {
    double g[9];
    rank2_recompose_v(f, V0, g, Vout);
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
        __builtin_amdgcn_sched_barrier(0);
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
        __builtin_amdgcn_sched_barrier(0);
    }
    return true;
}

// Lanes 3g, 3g+1, 3g+2 of a wave form solver group g (g < kGroups); a group splits the
// nine rows of a 9x9 eigen-solve's orthogonal basis three per lane while every
// lane of the group reduces the same matrix, which keeps both in registers.
constexpr int kGroups = 21;

__device__ __forceinline__ int solver_group() { return (int)(threadIdx.x & 63u) / 3; }
__device__ __forceinline__ int solver_part() {
    const int l = (int)(threadIdx.x & 63u);
    return l - 3 * (l / 3);
}

// Eigen-solve of the upper triangle `a` (45, overwritten) by a solver group:
// tridiagonalisation T = Q^T A Q (each lane accumulating three rows of Q),
// the eigenvalues of T by QL into d, the eigenvector x of the smallest one
// (first minimum) by inverse iteration on T, and Q x into vk in every lane
// of the group. All three lanes of the group must be active with the same
// `a`.
__device__ __forceinline__ int eigen_min_group(double* a, double* d, double* vk) {
    const int part = solver_part();
    const int base = 3 * solver_group();
    double q[3][9], e[9], d0[9], e0[9];
    tridiag9_rows<3>(a, 3 * part, q, d, e);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        d0[i] = d[i];
        e0[i] = e[i];
    }
    tql9_values(d, e);
    // first minimum (the value is tracked so that no array is indexed at run time)
    int k = 0;
    double dmin = d[0];
#pragma unroll
    for (int i = 1; i < 9; ++i) {
        if (d[i] < dmin) {
            dmin = d[i];
            k = i;
        }
    }
    double x[9];
    tinv9(d0, e0, dmin, x);
    // this lane's three rows of Q x
    double mine[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 9; ++j) acc = acc + q[i][j] * x[j];
        mine[i] = acc;
    }
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int i = 0; i < 3; ++i) vk[3 * t + i] = __shfl(mine[i], base + t);
    return k;
}

// The eigenvector of the smallest eigenvalue of a semidefinite 8-point normal
// matrix `a` (upper triangle, 45, overwritten) by a solver group, without the
// full eigen-solve (the host's eigen_min_inv9 step for step). Returns true if
// the system is degenerate (matching.rs:541-545), else false with the unit
// eigenvector in vk in every lane of the group:
//   T = Q^T A Q (each lane accumulating three rows of Q); degenerate when the
//   Sturm sequence of T - tau I, tau = 1e-10 G (G: Gershgorin bound of T),
//   has two or more negative terms; the eigenvector by inverse iteration on T
//   with shift 0 (partial-pivoting LU, pivots below 2^-52 G raised to it,
//   x_1 = unit(U^-1 1), x_{k+1} = unit((P L U)^-1 x_k)) until two iterates
//   agree to 2^-44 per component up to sign (at most 30 steps); vk = Q x.
// All three lanes of the group must be active with the same `a`.
__device__ __forceinline__ bool eigen_min_inv_group(double* a, double* vk) {
    const int part = solver_part();
    const int base = 3 * solver_group();
    double q[3][9], d[9], e[9];
    tridiag9_rows<3>(a, 3 * part, q, d, e);
    double g = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double rs = fabs(d[i]);
        if (i < 8) rs = rs + fabs(e[i]);
        if (i > 0) rs = rs + fabs(e[i - 1]);
        g = fmax(g, rs);
    }
    const double tau = 1e-10 * g;
    int neg = 0;
    double qs = d[0] - tau;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if (fabs(qs) < 0x1p-1000) qs = -0x1p-1000;
        if (qs < 0.0) ++neg;
        if (i < 8) qs = (d[i + 1] - tau) - e[i] * e[i] / qs;
    }
    const bool degenerate = neg >= 2;
    double x[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!degenerate) {
        const double eps3 = g * 0x1p-52;
        double dd[9], du[8], du2[7], dl[8], ui[9];
        bool swp[8];
#pragma unroll
        for (int i = 0; i < 9; ++i) dd[i] = d[i];
#pragma unroll
        for (int i = 0; i < 8; ++i) du[i] = dl[i] = e[i];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const bool keep = fabs(dd[i]) >= fabs(dl[i]);
            swp[i] = !keep;
            if (keep) {
                const double fact = dd[i] != 0.0 ? dl[i] / dd[i] : 0.0;
                dl[i] = fact;
                dd[i + 1] = dd[i + 1] - fact * du[i];
                if (i < 7) du2[i] = 0.0;
            } else {
                const double fact = dd[i] / dl[i];
                dd[i] = dl[i];
                dl[i] = fact;
                const double t = du[i];
                du[i] = dd[i + 1];
                dd[i + 1] = t - fact * dd[i + 1];
                if (i < 7) {
                    du2[i] = du[i + 1];
                    du[i + 1] = -fact * du[i + 1];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            double u = dd[i];
            if (fabs(u) < eps3) u = u >= 0.0 ? eps3 : -eps3;
            ui[i] = 1.0 / u;
        }
        double b[9], prev[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) b[i] = 1.0;
        for (int it = 0; it < 30; ++it) {
            if (it > 0) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if (swp[i]) {
                        const double t = b[i];
                        b[i] = b[i + 1];
                        b[i + 1] = t - dl[i] * b[i];
                    } else {
                        b[i + 1] = b[i + 1] - dl[i] * b[i];
                    }
                }
            }
            x[8] = b[8] * ui[8];
            x[7] = (b[7] - du[7] * x[8]) * ui[7];
#pragma unroll
            for (int i = 6; i >= 0; --i)
                x[i] = (b[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) * ui[i];
            unit9(x);
            bool conv = false;
            if (it > 0) {
                double dot = 0.0;
#pragma unroll
                for (int i = 0; i < 9; ++i) dot = dot + x[i] * prev[i];
                const double sg = dot >= 0.0 ? 1.0 : -1.0;
                double dm = 0.0;
#pragma unroll
                for (int i = 0; i < 9; ++i) dm = fmax(dm, fabs(x[i] - sg * prev[i]));
                conv = dm <= 0x1p-44;
            }
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                prev[i] = x[i];
                b[i] = x[i];
            }
            if (conv) break;
        }
    }
    // this lane's three rows of Q x
    double mine[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 9; ++j) acc = acc + q[i][j] * x[j];
        mine[i] = acc;
    }
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int i = 0; i < 3; ++i) vk[3 * t + i] = __shfl(mine[i], base + t);
    return degenerate;
}

// matching.rs:539-556 computed by a solver group (see eigen_min_group):
// the eigen-solve, the degeneracy test, the rank-2 projection and the
// denormalisation. ata is overwritten.
__device__ bool solve_8pt_group(double* ata, const double* t1, const double* t2, double* F) {
    double fn[9];
    if (eigen_min_inv_group(ata, fn)) return false;
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
    // branch-free: the quotient is discarded when den <= 0
    const double err = e * e / den;
    return den <= 0.0 ? (double)INFINITY : err;
}

}  // namespace rdev
}  // namespace skh
