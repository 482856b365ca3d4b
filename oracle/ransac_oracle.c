/*
 * CPU oracle, fundamental-matrix RANSAC half. TEST INFRASTRUCTURE ONLY (see
 * oracle.h): the checker for the HIP path, never the product.
 *
 * Restates crates/core/src/matching.rs:217-1104 and
 * crates/core/src/two_view.rs:605-642 (homography_dlt). The reference's
 * linear algebra comes from nalgebra 0.33.3 (Cargo.lock:166-168):
 * SymmetricEigen (matching.rs:539, two_view.rs:638), SVD (matching.rs:550,
 * 577), Cholesky (:714), try_inverse (:263, :340, :359, :675). nalgebra is not
 * vendored and cannot be built here, so the oracle uses its own solvers with
 * the same mathematical contract:
 *   - 8-point solve (matching.rs:539-546): Householder tridiagonalisation,
 *     the degeneracy test by a Sturm count, the smallest eigenvector by
 *     inverse iteration with shift 0 (no eigen-solve);
 *   - symmetric 9x9 eigen for the DLT refits: Householder
 *     tridiagonalisation, the eigenvalues by implicit QL (nalgebra's
 *     SymmetricEigen also reduces to tridiagonal form and iterates with
 *     shifts), the smallest one's eigenvector by inverse iteration;
 *   - homography of a four-point sample (two_view.rs:607-642 on 4 points):
 *     the exact null vector of its DLT system in closed form (projective
 *     basis construction) instead of the A^T A eigenvector;
 *   - 3x3 SVD for the rank-2 projection: one-sided Jacobi (Hestenes 1958),
 *     inside the Sampson LM started from the current F's singular basis;
 *   - Cholesky solve of the LM step: as LDL^T (no square roots),
 *     right-looking column factorisation, forward/back solves;
 *   - 3x3 inverse: cofactor expansion.
 * The GPU performs the same operations in the same order and produces the
 * same bits. The F bits therefore pin the HIP path to this oracle, not to
 * the Rust binary (parity with the reference is at the level of its tests).
 *
 * Matrices are row-major double[9] / double[81]. All sums run in the
 * reference's order (first element first).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---------------------------------------------------------------- helpers */

static const double NM[6] = {3.0, 2.0, 1.5, 1.0, 1.0, 1.0}; /* matching.rs:998, :1043 */

static inline int all_finite2(const double* p, size_t n) {
    for (size_t i = 0; i < 2 * n; ++i)
        if (!isfinite(p[i])) return 0;
    return 1;
}

/* y = M x (nalgebra gemv order: column 0, 1, 2). */
static inline void mv3(const double* m, const double* x, double* y) {
    double a = m[0] * x[0] + m[1] * x[1] + m[2] * x[2];
    double b = m[3] * x[0] + m[4] * x[1] + m[5] * x[2];
    double c = m[6] * x[0] + m[7] * x[1] + m[8] * x[2];
    y[0] = a; y[1] = b; y[2] = c;
}

/* y = M^T x. */
static inline void mtv3(const double* m, const double* x, double* y) {
    double a = m[0] * x[0] + m[3] * x[1] + m[6] * x[2];
    double b = m[1] * x[0] + m[4] * x[1] + m[7] * x[2];
    double c = m[2] * x[0] + m[5] * x[1] + m[8] * x[2];
    y[0] = a; y[1] = b; y[2] = c;
}

/* C = A B. */
static inline void mm3(const double* A, const double* B, double* C) {
    double t[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            t[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
    memcpy(C, t, sizeof t);
}

static inline void tr3(const double* A, double* T) {
    double t[9] = {A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8]};
    memcpy(T, t, sizeof t);
}

/* Frobenius norm, summed column by column (nalgebra norm_squared). */
static inline double fro3(const double* m) {
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
        double col = m[c] * m[c] + m[3 + c] * m[3 + c] + m[6 + c] * m[6 + c];
        s = s + col;
    }
    return sqrt(s);
}

static inline double norm2(double x, double y) { return sqrt(x * x + y * y); }

/* 3x3 inverse by cofactors; returns 0 when the determinant is zero. */
static int inv3(const double* m, double* out) {
    double m11 = m[4] * m[8] - m[5] * m[7];
    double m12 = m[3] * m[8] - m[5] * m[6];
    double m13 = m[3] * m[7] - m[4] * m[6];
    double det = m[0] * m11 - m[1] * m12 + m[2] * m13;
    if (det == 0.0) return 0;
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
    memcpy(out, o, sizeof o);
    return 1;
}

/* LCG step of the samplers (matching.rs:957-962, :315-320, :884-889). */
static inline uint64_t rnd(uint64_t* st, uint64_t m) {
    *st = *st * 6364136223846793005ULL + 1442695040888963407ULL;
    return (*st >> 33) % m;
}

/* ------------------------------------------------------------ eigen / SVD */

/* sqrt(a^2 + b^2), formed directly: the operands (shifts and couplings of
 * the tridiagonal form of a Hartley-normalised 9x9 A^T A, entries of order
 * the point count) stay many decades away from overflow, and a sum that
 * underflows to 0 is the r == 0 case of the sweep. */
static inline double hyp2(double a, double b) { return sqrt(a * a + b * b); }

/* Householder tridiagonalisation of the symmetric 9x9 `a` (row-major, full,
 * overwritten; Golub & Van Loan, Matrix Computations 4th ed., Alg. 8.3.1),
 * accumulating q = H_0 H_1 ... H_6 row by row. Step k reflects
 * x = a[k+1..8][k] onto beta e_1 with beta = -sign(x_0) |x| and
 * v = (1, x_1 / (x_0 - beta), ...), tau = (beta - x_0) / beta; the trailing
 * block B becomes B - v w^T - w v^T with p = tau B v,
 * w = p - (tau / 2)(v^T p) v. d: diagonal, e[i]: the (i, i+1) coupling,
 * e[8] = 0. */
static void tridiag9(double* a, double* q, double* d, double* e) {
    for (int i = 0; i < 81; ++i) q[i] = (i % 10 == 0) ? 1.0 : 0.0;
    for (int k = 0; k < 7; ++k) {
        double xn2 = 0.0;
        for (int i = k + 2; i < 9; ++i) xn2 = xn2 + a[9 * i + k] * a[9 * i + k];
        const double x0 = a[9 * (k + 1) + k];
        if (xn2 == 0.0) continue;
        const double nrm = sqrt(x0 * x0 + xn2);
        const double beta = x0 >= 0.0 ? -nrm : nrm;
        const double tau = (beta - x0) / beta;
        const double rs = 1.0 / (x0 - beta);
        double v[9], p[9], w[9];
        v[k + 1] = 1.0;
        for (int i = k + 2; i < 9; ++i) v[i] = a[9 * i + k] * rs;
        for (int i = k + 1; i < 9; ++i) {
            double sum = 0.0;
            for (int j = k + 1; j < 9; ++j) sum = sum + a[9 * i + j] * v[j];
            p[i] = tau * sum;
        }
        double vp = 0.0;
        for (int i = k + 1; i < 9; ++i) vp = vp + v[i] * p[i];
        const double kk = 0.5 * tau * vp;
        for (int i = k + 1; i < 9; ++i) w[i] = p[i] - kk * v[i];
        for (int i = k + 1; i < 9; ++i)
            for (int j = i; j < 9; ++j) {
                const double nv = a[9 * i + j] - (v[i] * w[j] + w[i] * v[j]);
                a[9 * i + j] = nv;
                a[9 * j + i] = nv;
            }
        a[9 * (k + 1) + k] = beta;
        a[9 * k + (k + 1)] = beta;
        for (int i = k + 2; i < 9; ++i) {
            a[9 * i + k] = 0.0;
            a[9 * k + i] = 0.0;
        }
        for (int r = 0; r < 9; ++r) {
            double sum = 0.0;
            for (int j = k + 1; j < 9; ++j) sum = sum + q[9 * r + j] * v[j];
            const double t = tau * sum;
            for (int j = k + 1; j < 9; ++j) q[9 * r + j] = q[9 * r + j] - t * v[j];
        }
    }
    for (int i = 0; i < 9; ++i) d[i] = a[9 * i + i];
    for (int i = 0; i < 8; ++i) e[i] = a[9 * (i + 1) + i];
    e[8] = 0.0;
}

#define QL_MAX_ITERS 30

/* Eigenvalues of the symmetric tridiagonal (d, e) by implicit QL with shifts
 * (Wilkinson & Reinsch, Handbook for Automatic Computation II, tql1: the
 * eigenvalue form of tql2). Each rotation's radius is r = sqrt(f^2 + g^2),
 * its cosine and sine are g and f times 1 / r. Eigenvalue l is split off
 * once e[l] is negligible against its neighbours; at most QL_MAX_ITERS
 * sweeps per eigenvalue. */
static void tql9_values(double* d, double* e) {
    for (int l = 0; l < 9; ++l) {
        for (int iter = 0; iter < QL_MAX_ITERS; ++iter) {
            int m = l;
            for (; m < 8; ++m) {
                const double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) + dd == dd) break;
            }
            if (m == l) break;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hyp2(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int underflow = 0;
            for (int i = m - 1; i >= l; --i) {
                const double f = s * e[i], b = c * e[i];
                r = hyp2(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] = d[i + 1] - p;
                    e[m] = 0.0;
                    underflow = 1;
                    break;
                }
                const double ri = 1.0 / r;
                s = f * ri;
                c = g * ri;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
            }
            if (underflow) continue;
            d[l] = d[l] - p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
}

/* x / max|x_i|, then divided by its 2-norm (sums in index order). */
static void unit9(double* x) {
    double mx = 0.0;
    for (int i = 0; i < 9; ++i) mx = fmax(mx, fabs(x[i]));
    const double sm = 1.0 / mx;
    double ss = 0.0;
    for (int i = 0; i < 9; ++i) {
        x[i] = x[i] * sm;
        ss = ss + x[i] * x[i];
    }
    const double sn = 1.0 / sqrt(ss);
    for (int i = 0; i < 9; ++i) x[i] = x[i] * sn;
}

/* Unit eigenvector of the symmetric tridiagonal (d, e) for its eigenvalue
 * lam by inverse iteration (Wilkinson, The Algebraic Eigenvalue Problem,
 * 1965, ch. 5.54; EISPACK tinvit): T - lam I = P L U by Gaussian
 * elimination with partial pivoting (LAPACK dgttrf's layout: U has two
 * superdiagonals), every pivot of U smaller than eps3 = 2^-52 |T| in
 * magnitude replaced by eps3 with its sign (|T| = the largest absolute row
 * sum); then x1 = U^-1 (1, ..., 1) (the start vector b with L^-1 P b = 1),
 * x2 = (P L U)^-1 unit(x1), x = unit(x2). */
static void tinv9(const double* d, const double* e, double lam, double* x) {
    double tn = 0.0;
    for (int i = 0; i < 9; ++i) {
        double rs = fabs(d[i]);
        if (i < 8) rs = rs + fabs(e[i]);
        if (i > 0) rs = rs + fabs(e[i - 1]);
        tn = fmax(tn, rs);
    }
    const double eps3 = tn * 0x1p-52;
    double dd[9], du[8], du2[7], dl[8];
    int swp[8];
    for (int i = 0; i < 9; ++i) dd[i] = d[i] - lam;
    for (int i = 0; i < 8; ++i) du[i] = dl[i] = e[i];
    for (int i = 0; i < 8; ++i) {
        if (fabs(dd[i]) >= fabs(dl[i])) {
            swp[i] = 0;
            const double fact = dd[i] != 0.0 ? dl[i] / dd[i] : 0.0;
            dl[i] = fact;
            dd[i + 1] = dd[i + 1] - fact * du[i];
            if (i < 7) du2[i] = 0.0;
        } else {
            swp[i] = 1;
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
    for (int i = 0; i < 9; ++i) {
        double u = dd[i];
        if (fabs(u) < eps3) u = u >= 0.0 ? eps3 : -eps3;
        ui[i] = 1.0 / u;
    }
    double b[9];
    for (int i = 0; i < 9; ++i) b[i] = 1.0;
    for (int it = 0; it < 2; ++it) {
        if (it == 1)
            for (int i = 0; i < 8; ++i) {
                if (swp[i]) {
                    const double t = b[i];
                    b[i] = b[i + 1];
                    b[i + 1] = t - dl[i] * b[i];
                } else {
                    b[i + 1] = b[i + 1] - dl[i] * b[i];
                }
            }
        x[8] = b[8] * ui[8];
        x[7] = (b[7] - du[7] * x[8]) * ui[7];
        for (int i = 6; i >= 0; --i) x[i] = (b[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) * ui[i];
        unit9(x);
        memcpy(b, x, sizeof b);
    }
}

/* The eigenvector of the smallest eigenvalue of the symmetric positive
 * semidefinite 9x9 A (an 8-point normal matrix) without the full eigen-solve.
 * Returns 1 if the system is degenerate (fundamental_8pt's test,
 * matching.rs:541-545: the second smallest eigenvalue at most 1e-10 of the
 * largest), else 0 with a unit eigenvector in vec.
 *   T = Q^T A Q (tridiag9). Degeneracy: the eigenvalues of T below
 *   tau = 1e-10 G, G = max_i |d_i| + |e_{i-1}| + |e_i| >= lambda_max
 *   (Gershgorin), counted by the Sturm sequence q_0 = d_0 - tau,
 *   q_i = (d_i - tau) - e_{i-1}^2 / q_{i-1} (|q_i| < 2^-1000 taken as
 *   -2^-1000): two or more negative q_i. Against lambda_max itself this also
 *   rejects the systems with lambda_2 / lambda_max in (1e-10, 1e-10 G /
 *   lambda_max], where G / lambda_max <= 3.
 *   Eigenvector: inverse iteration on T with shift 0 (the smallest eigenvalue
 *   of a semidefinite matrix is the one nearest 0): T = P L U as in tinv9,
 *   x_1 = unit(U^-1 (1, ..., 1)), x_{k+1} = unit((P L U)^-1 x_k) until two
 *   successive iterates agree to 2^-44 in every component (up to sign), at
 *   most 30 steps; vec = Q x. The error shrinks by lambda_1 / lambda_2 per
 *   step: one or two steps for an exact 8-point sample, a few for a
 *   least-squares refit. */
#define INV_MAX_STEPS 30
static int eigen_min_inv9(const double* A, double* vec) {
    double a[81], q[81], d[9], e[9];
    memcpy(a, A, sizeof a);
    tridiag9(a, q, d, e);
    double g = 0.0;
    for (int i = 0; i < 9; ++i) {
        double rs = fabs(d[i]);
        if (i < 8) rs = rs + fabs(e[i]);
        if (i > 0) rs = rs + fabs(e[i - 1]);
        g = fmax(g, rs);
    }
    const double tau = 1e-10 * g;
    int neg = 0;
    double qs = d[0] - tau;
    for (int i = 0;; ++i) {
        if (fabs(qs) < 0x1p-1000) qs = -0x1p-1000;
        if (qs < 0.0) ++neg;
        if (i == 8) break;
        qs = (d[i + 1] - tau) - e[i] * e[i] / qs;
    }
    if (neg >= 2) return 1;
    /* T = P L U with partial pivoting (tinv9's elimination, shift 0) */
    const double eps3 = g * 0x1p-52;
    double dd[9], du[8], du2[7], dl[8], ui[9];
    int swp[8];
    for (int i = 0; i < 9; ++i) dd[i] = d[i];
    for (int i = 0; i < 8; ++i) du[i] = dl[i] = e[i];
    for (int i = 0; i < 8; ++i) {
        if (fabs(dd[i]) >= fabs(dl[i])) {
            swp[i] = 0;
            const double fact = dd[i] != 0.0 ? dl[i] / dd[i] : 0.0;
            dl[i] = fact;
            dd[i + 1] = dd[i + 1] - fact * du[i];
            if (i < 7) du2[i] = 0.0;
        } else {
            swp[i] = 1;
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
    for (int i = 0; i < 9; ++i) {
        double u = dd[i];
        if (fabs(u) < eps3) u = u >= 0.0 ? eps3 : -eps3;
        ui[i] = 1.0 / u;
    }
    double b[9], x[9];
    for (int i = 0; i < 9; ++i) b[i] = 1.0;
    for (int it = 0; it < INV_MAX_STEPS; ++it) {
        if (it > 0)
            for (int i = 0; i < 8; ++i) {
                if (swp[i]) {
                    const double t = b[i];
                    b[i] = b[i + 1];
                    b[i + 1] = t - dl[i] * b[i];
                } else {
                    b[i + 1] = b[i + 1] - dl[i] * b[i];
                }
            }
        x[8] = b[8] * ui[8];
        x[7] = (b[7] - du[7] * x[8]) * ui[7];
        for (int i = 6; i >= 0; --i) x[i] = (b[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) * ui[i];
        unit9(x);
        int conv = 0;
        if (it > 0) {
            double dot = 0.0;
            for (int i = 0; i < 9; ++i) dot = dot + x[i] * vec[i];
            const double sg = dot >= 0.0 ? 1.0 : -1.0;
            double dm = 0.0;
            for (int i = 0; i < 9; ++i) dm = fmax(dm, fabs(x[i] - sg * vec[i]));
            conv = dm <= 0x1p-44;
        }
        memcpy(vec, x, sizeof x); /* vec: the previous unit iterate */
        memcpy(b, x, sizeof b);
        if (conv) break;
    }
    double y[9];
    for (int i = 0; i < 9; ++i) {
        double s = 0.0;
        for (int j = 0; j < 9; ++j) s = s + q[9 * i + j] * vec[j];
        y[i] = s;
    }
    memcpy(vec, y, sizeof y);
    return 0;
}

/* The eigenvalues of the symmetric 9x9 A (ev, in the order the QL sweep
 * leaves them) and a unit eigenvector of the first smallest one (vec):
 * tridiagonalisation T = Q^T A Q, eigenvalues of T by QL, the eigenvector
 * of T by inverse iteration, vec = Q x. */
static int imin9(const double* e);
static void eigen_min9(const double* A, double* ev, double* vec) {
    double a[81], q[81], d[9], e[9], d0[9], e0[9], x[9];
    memcpy(a, A, sizeof a);
    tridiag9(a, q, d, e);
    memcpy(d0, d, sizeof d0);
    memcpy(e0, e, sizeof e0);
    tql9_values(d, e);
    tinv9(d0, e0, d[imin9(d)], x);
    for (int i = 0; i < 9; ++i) {
        double s = 0.0;
        for (int j = 0; j < 9; ++j) s = s + q[9 * i + j] * x[j];
        vec[i] = s;
    }
    memcpy(ev, d, sizeof d);
}

void orc_eigen_min9(const double* A, double* evals, double* evec) { eigen_min9(A, evals, evec); }
int orc_eigen_min_inv9(const double* A, double* evec) { return eigen_min_inv9(A, evec); }

static int imin9(const double* e) {
    int k = 0;
    for (int i = 1; i < 9; ++i)
        if (e[i] < e[k]) k = i;
    return k;
}

static int cmp_double(const void* x, const void* y) {
    double a = *(const double*)x, b = *(const double*)y;
    return (a > b) - (a < b);
}

/* One-sided Jacobi SVD of F (3x3): W = F V with orthogonal columns, then the
 * column of least norm is dropped and F' = W' V^T (matching.rs:550-553,
 * :577-580: zero the smallest singular value, recompose). The rotations start
 * from the identity (V0 = NULL: W = F) or from an orthogonal basis V0
 * (row-major, column j = basis vector j: W = F V0, V = V0), e.g. the basis of
 * a nearby matrix, which leaves fewer sweeps; Vout (if set) receives V. A
 * column pair is rotated while gamma^2 > 1e-30 alpha beta (|gamma| above
 * 1e-15 of the geometric mean of the two squared column norms). */
#define SVD3_MAX_SWEEPS 30
static void rank2_recompose_v(const double* F, const double* V0, double* out, double* Vout) {
    double w[9], v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (V0) {
        for (int r = 0; r < 3; ++r)
            for (int j = 0; j < 3; ++j)
                w[3 * r + j] = F[3 * r] * V0[j] + F[3 * r + 1] * V0[3 + j] + F[3 * r + 2] * V0[6 + j];
        memcpy(v, V0, sizeof v);
    } else {
        memcpy(w, F, sizeof w);
    }
    static const int PP[3] = {0, 0, 1}, QQ[3] = {1, 2, 2};
    for (int sweep = 0; sweep < SVD3_MAX_SWEEPS; ++sweep) {
        int rotated = 0;
        for (int r = 0; r < 3; ++r) {
            int p = PP[r], q = QQ[r];
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
                rotated = 1;
            }
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
    if (Vout) memcpy(Vout, v, sizeof v);
    double sig[3];
    for (int c = 0; c < 3; ++c) sig[c] = w[c] * w[c] + w[3 + c] * w[3 + c] + w[6 + c] * w[6 + c];
    int m = 0;
    for (int c = 1; c < 3; ++c)
        if (sig[c] < sig[m]) m = c;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
            for (int i = 0; i < 3; ++i) {
                if (i == m) continue;
                acc = acc + w[3 * r + i] * v[3 * c + i];
            }
            out[3 * r + c] = acc;
        }
}

static void rank2_recompose(const double* F, double* out) { rank2_recompose_v(F, NULL, out, NULL); }

/* matching.rs:576-583 (V0 / Vout as in rank2_recompose_v). */
static int rank2_unit_v(const double* f, const double* V0, double* out, double* Vout) {
    double g[9];
    rank2_recompose_v(f, V0, g, Vout);
    double n = fro3(g);
    if (!(isfinite(n) && n > 0.0)) return 0;
    for (int i = 0; i < 9; ++i) out[i] = g[i] / n;
    return 1;
}


/* --------------------------------------------------------------- geometry */

/* matching.rs:225-241. */
static void axis_spread(const double* p, size_t len, double* lo_out, double* hi_out) {
    double n = (double)(len > 0 ? len : 1);
    double cx = 0.0, cy = 0.0;
    for (size_t i = 0; i < len; ++i) { cx = cx + p[2 * i]; cy = cy + p[2 * i + 1]; }
    cx = cx / n; cy = cy / n;
    double sxx = 0.0, sxy = 0.0, syy = 0.0;
    for (size_t i = 0; i < len; ++i) {
        double dx = p[2 * i] - cx, dy = p[2 * i + 1] - cy;
        sxx += dx * dx;
        sxy += dx * dy;
        syy += dy * dy;
    }
    sxx = sxx / n; sxy = sxy / n; syy = syy / n;
    double tr = sxx + syy;
    double dd = sxx - syy;
    double disc = sqrt(dd * dd + 4.0 * sxy * sxy);
    double hi = 0.5 * (tr + disc);
    double lo = 0.5 * (tr - disc);
    if (!(lo >= 0.0)) lo = 0.0; /* f64::max(0.0) */
    *lo_out = sqrt(lo);
    *hi_out = sqrt(hi);
}

/* matching.rs:244-247. */
static int nearly_collinear(const double* p, size_t n) {
    double lo, hi;
    axis_spread(p, n, &lo, &hi);
    return !(hi > 0.0 && lo > 0.02 * hi);
}

/* matching.rs:499-505. */
void orc_normalizer(const double* p, size_t len, double* T) {
    double n = (double)len;
    double cx = 0.0, cy = 0.0;
    for (size_t i = 0; i < len; ++i) { cx = cx + p[2 * i]; cy = cy + p[2 * i + 1]; }
    cx = cx / n; cy = cy / n;
    double d = 0.0;
    for (size_t i = 0; i < len; ++i) d = d + norm2(p[2 * i] - cx, p[2 * i + 1] - cy);
    d = d / n;
    double s = d > 0.0 ? 1.4142135623730951 / d : 1.0;
    double t[9] = {s, 0.0, -s * cx, 0.0, s, -s * cy, 0.0, 0.0, 1.0};
    memcpy(T, t, sizeof t);
}

/* matching.rs:512-557. */
int orc_fundamental_8pt(const double* x1, const double* x2, size_t n, double* F) {
    if (n < 8 || !all_finite2(x1, n) || !all_finite2(x2, n)) return 0;
    if (nearly_collinear(x1, n) || nearly_collinear(x2, n)) return 0;
    double t1[9], t2[9];
    orc_normalizer(x1, n, t1);
    orc_normalizer(x2, n, t2);
    double ata[81];
    memset(ata, 0, sizeof ata);
    for (size_t i = 0; i < n; ++i) {
        double pa[3] = {x1[2 * i], x1[2 * i + 1], 1.0}, pb[3] = {x2[2 * i], x2[2 * i + 1], 1.0};
        double a[3], b[3];
        mv3(t1, pa, a);
        mv3(t2, pb, b);
        double row[9] = {b[0] * a[0], b[0] * a[1], b[0], b[1] * a[0], b[1] * a[1], b[1],
                         a[0], a[1], 1.0};
        for (int r = 0; r < 9; ++r)
            for (int c = 0; c < 9; ++c) ata[9 * r + c] = ata[9 * r + c] + row[r] * row[c];
    }
    double vec[9];
    if (eigen_min_inv9(ata, vec)) return 0;
    double fn[9];
    for (int i = 0; i < 9; ++i) fn[i] = vec[i];
    double r2[9];
    rank2_recompose(fn, r2);
    double t2t[9], tmp[9], f[9];
    tr3(t2, t2t);
    mm3(t2t, r2, tmp);
    mm3(tmp, t1, f);
    double nn = fro3(f);
    if (!(isfinite(nn) && nn > 0.0)) return 0;
    for (int i = 0; i < 9; ++i) F[i] = f[i] / nn;
    return 1;
}

/* matching.rs:561-573. */
double orc_sampson_error(const double* F, const double* p, const double* q) {
    double a[3] = {p[0], p[1], 1.0}, b[3] = {q[0], q[1], 1.0};
    double fa[3], ftb[3];
    mv3(F, a, fa);
    mtv3(F, b, ftb);
    double e = b[0] * fa[0] + b[1] * fa[1] + b[2] * fa[2];
    double den = fa[0] * fa[0] + fa[1] * fa[1] + ftb[0] * ftb[0] + ftb[1] * ftb[1];
    if (den <= 0.0) return INFINITY;
    return e * e / den;
}

/* matching.rs:591-618. */
void orc_sampson_residual_jacobian(const double* g, double s1, double s2, const double* a,
                                   const double* b, double* r, double* j) {
    double ga[3], gtb[3];
    mv3(g, a, ga);
    mtv3(g, b, gtb);
    double e = b[0] * ga[0] + b[1] * ga[1] + b[2] * ga[2];
    double q1 = s1 * s1, q2 = s2 * s2;
    double den = q2 * (ga[0] * ga[0] + ga[1] * ga[1]) + q1 * (gtb[0] * gtb[0] + gtb[1] * gtb[1]);
    if (den <= 0.0) {
        *r = 0.0;
        for (int k = 0; k < 9; ++k) j[k] = 0.0;
        return;
    }
    double inv = 1.0 / sqrt(den);
    double c = e * inv * inv * inv;
    double u[3] = {q2 * ga[0], q2 * ga[1], 0.0};
    double v[3] = {q1 * gtb[0], q1 * gtb[1], 0.0};
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
            j[3 * k + l] = b[k] * a[l] * inv - c * (u[k] * a[l] + b[k] * v[l]);
    *r = e * inv;
}

/* matching.rs:621-641, summed into the cost sum_e e^2 in order; each
 * squared residual e^2 = (b . G a)^2 / den is formed without the square
 * root of the residual itself. */
static double sampson_cost(const double* g, double s1, double s2, const double* a,
                           const double* b, size_t n) {
    double q1 = s1 * s1, q2 = s2 * s2;
    double cost = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double ga[3], gtb[3];
        mv3(g, a + 3 * i, ga);
        mtv3(g, b + 3 * i, gtb);
        double den = q2 * (ga[0] * ga[0] + ga[1] * ga[1]) + q1 * (gtb[0] * gtb[0] + gtb[1] * gtb[1]);
        double e2 = 0.0;
        if (den > 0.0) {
            const double* bb = b + 3 * i;
            const double num = bb[0] * ga[0] + bb[1] * ga[1] + bb[2] * ga[2];
            e2 = num * num / den;
        }
        cost = cost + e2;
    }
    return cost;
}

/* Solve m x = rhs for the symmetric 9x9 m of an LM step (matching.rs:714,
 * nalgebra's Cholesky there) by its LDL^T factorisation, column by column
 * with right-looking updates: L unit lower, D diagonal (the Cholesky factor
 * is L D^1/2, so no square root is taken); then L y = rhs, y := D^-1 y
 * (through the reciprocals of D), L^T x = y. Returns 0 if a pivot D_j is
 * not positive: m is not positive definite, where the reference's Cholesky
 * fails. */
static int cholesky_solve9(const double* m_in, const double* rhs, double* x) {
    double A[81], D[9], rd[9];
    memcpy(A, m_in, sizeof A);
    for (int j = 0; j < 9; ++j) {
        for (int k = 0; k < j; ++k) {
            const double w = A[9 * j + k] * D[k];
            for (int i = j; i < 9; ++i) A[9 * i + j] = A[9 * i + j] - w * A[9 * i + k];
        }
        D[j] = A[9 * j + j];
        if (!(D[j] > 0.0)) return 0;
        rd[j] = 1.0 / D[j];
        for (int i = j + 1; i < 9; ++i) A[9 * i + j] = A[9 * i + j] * rd[j];
    }
    double y[9];
    memcpy(y, rhs, sizeof y);
    for (int i = 0; i < 9; ++i)
        for (int k = i + 1; k < 9; ++k) y[k] = y[k] - A[9 * k + i] * y[i];
    for (int i = 0; i < 9; ++i) y[i] = y[i] * rd[i];
    for (int i = 8; i >= 0; --i) {
        double dot = 0.0;
        for (int k = i + 1; k < 9; ++k) dot = dot + A[9 * k + i] * y[k];
        y[i] = y[i] - dot;
    }
    memcpy(x, y, sizeof y);
    return 1;
}

static __thread uint64_t g_lm_calls, g_lm_conv, g_lm_stall0, g_iters, g_lo_runs;
static __thread uint64_t g_lm_iters, g_lm_evals; /* LM iterations and cost evaluations */

/* matching.rs:665-760. stop: 0 = converged, 1 = stalled, 2 = cap. */
int orc_refine_sampson(const double* F, const double* x1, const double* x2, size_t n,
                       size_t iters, double* out, uint64_t* taken_out, int* stop_out) {
    memcpy(out, F, sizeof(double) * 9);
    if (taken_out) *taken_out = 0;
    if (stop_out) *stop_out = 2;
    if (n < 8) return 0;
    double t1[9], t2[9], t1i[9], t2i[9];
    orc_normalizer(x1, n, t1);
    orc_normalizer(x2, n, t2);
    if (!inv3(t1, t1i) || !inv3(t2, t2i)) return 0;
    double s1 = t1[0], s2 = t2[0];
    double* a = (double*)malloc(sizeof(double) * 3 * n);
    double* b = (double*)malloc(sizeof(double) * 3 * n);
    for (size_t i = 0; i < n; ++i) {
        double pa[3] = {x1[2 * i], x1[2 * i + 1], 1.0}, pb[3] = {x2[2 * i], x2[2 * i + 1], 1.0};
        mv3(t1, pa, a + 3 * i);
        mv3(t2, pb, b + 3 * i);
    }
    double t2it[9], tmp[9], g0[9], g[9];
    tr3(t2i, t2it);
    mm3(t2it, F, tmp);
    mm3(tmp, t1i, g0);
    /* V: the right singular basis of g; each damped step's rank-2 projection starts its
     * Jacobi rotations from it (g + d is close to g) */
    double V[9];
    if (!rank2_unit_v(g0, NULL, g, V)) { free(a); free(b); return 0; }
    double cost = sampson_cost(g, s1, s2, a, b, n);
    double lambda = 1e-3;
    uint64_t taken = 0;
    int stop = 2;
    for (size_t it = 0; it < iters; ++it) {
        g_lm_iters++;
        double jtj[81], jtr[9];
        memset(jtj, 0, sizeof jtj);
        memset(jtr, 0, sizeof jtr);
        for (size_t i = 0; i < n; ++i) {
            double ri, ji[9];
            orc_sampson_residual_jacobian(g, s1, s2, a + 3 * i, b + 3 * i, &ri, ji);
            for (int c = 0; c < 9; ++c)
                for (int r = c; r < 9; ++r) jtj[9 * r + c] = ji[c] * ji[r] + jtj[9 * r + c];
            for (int k = 0; k < 9; ++k) jtr[k] = jtr[k] + ji[k] * ri;
        }
        for (int r = 0; r < 9; ++r)
            for (int c = r + 1; c < 9; ++c) jtj[9 * r + c] = jtj[9 * c + r];
        int improved = 0, accepted = 0;
        for (int tries = 0; tries < 8; ++tries) {
            double m[81];
            memcpy(m, jtj, sizeof m);
            for (int k = 0; k < 9; ++k) m[9 * k + k] = m[9 * k + k] + lambda * (jtj[9 * k + k] + 1e-12);
            double nrhs[9], d[9];
            for (int k = 0; k < 9; ++k) nrhs[k] = -jtr[k];
            if (!cholesky_solve9(m, nrhs, d)) { lambda *= 10.0; continue; }
            double gs[9], cand[9];
            for (int k = 0; k < 9; ++k) gs[k] = g[k] + d[k];
            double Vc[9];
            if (!rank2_unit_v(gs, V, cand, Vc)) { lambda *= 10.0; continue; }
            double c = sampson_cost(cand, s1, s2, a, b, n);
            g_lm_evals++;
            if (c < cost) {
                memcpy(g, cand, sizeof g);
                memcpy(V, Vc, sizeof V);
                double rel = (cost - c) / (cost > 1e-300 ? cost : 1e-300);
                cost = c;
                lambda = lambda * 0.1 > 1e-9 ? lambda * 0.1 : 1e-9;
                taken += 1;
                accepted = 1;
                improved = rel > 1e-10;
                break;
            }
            lambda *= 10.0;
        }
        if (!improved) {
            stop = accepted ? 0 : 1;
            break;
        }
    }
    free(a);
    free(b);
    g_lm_calls++;
    if (stop == 0) g_lm_conv++;
    if (stop == 1 && taken == 0) g_lm_stall0++;
    if (taken_out) *taken_out = taken;
    if (stop_out) *stop_out = stop;
    double t2t[9], fp[9];
    tr3(t2, t2t);
    mm3(t2t, g, tmp);
    mm3(tmp, t1, fp);
    double nn = fro3(fp);
    if (isfinite(nn) && nn > 0.0)
        for (int i = 0; i < 9; ++i) out[i] = fp[i] / nn;
    return 1;
}

/* Binary exponentiation (Rust powi, compiler-rt __powidf2). */
static double powi_(double a, int b) {
    int recip = b < 0;
    unsigned ub = recip ? (unsigned)(-b) : (unsigned)b;
    double r = 1.0;
    for (;;) {
        if (ub & 1) r *= a;
        ub /= 2;
        if (ub == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

/* matching.rs:799-811. */
uint64_t orc_adaptive_iterations(double w, int32_t sample, double confidence, uint64_t max_iters) {
    uint64_t floor_ = 50 < max_iters ? 50 : max_iters;
    double wc = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
    double p_good = powi_(wc, sample);
    if (p_good >= 1.0) return floor_;
    double denom = log1p(-p_good);
    const double cmax = 1.0 - 1e-15;
    double cc = confidence < 0.0 ? 0.0 : (confidence > cmax ? cmax : confidence);
    double need = log1p(-cc) / denom;
    if (!isfinite(need) || need >= (double)max_iters) return max_iters;
    double ce = ceil(need);
    uint64_t v = ce <= 0.0 ? 0 : (uint64_t)ce;
    if (v < floor_) v = floor_;
    if (v > max_iters) v = max_iters;
    return v;
}

/* matching.rs:825-840. */
double orc_epipolar_band_probability(const double* x2, size_t n, double th_px) {
    double lox = INFINITY, loy = INFINITY, hix = -INFINITY, hiy = -INFINITY;
    for (size_t i = 0; i < n; ++i) {
        lox = fmin(lox, x2[2 * i]); loy = fmin(loy, x2[2 * i + 1]);
        hix = fmax(hix, x2[2 * i]); hiy = fmax(hiy, x2[2 * i + 1]);
    }
    double dx = hix - lox, dy = hiy - loy;
    double area = dx * dy;
    if (!(isfinite(area) && area > 0.0)) return 1.0;
    double v = 2.0 * th_px * norm2(dx, dy) / area;
    return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}

/* matching.rs:843-865. */
double orc_binomial_tail(uint64_t n, uint64_t k, double p) {
    if (k == 0) return 1.0;
    if (k > n || p <= 0.0) return 0.0;
    if (p >= 1.0) return 1.0;
    double lp = log(p), lq = log1p(-p);
    double log_c = 0.0, sum = 0.0;
    for (uint64_t i = 0; i <= n; ++i) {
        if (i > 0) log_c += log((double)(n - i + 1)) - log((double)i);
        if (i >= k) sum += exp(log_c + (double)i * lp + (double)(n - i) * lq);
    }
    return sum < 1.0 ? sum : 1.0;
}

/* matching.rs:869-874 (FREE_FIT = 8, :821). */
double orc_chance_inlier_probability(uint64_t n, uint64_t cnt, double p) {
    if (cnt <= 8) return 1.0;
    return orc_binomial_tail(n > 8 ? n - 8 : 0, cnt - 8, p);
}

/* matching.rs:878-906. */
static size_t off_line_count(const double* p, size_t n, double th_px) {
    if (n < 3) return 0;
    uint64_t st = 0x6A09E667F3BCC909ULL;
    size_t best = 0;
    for (int it = 0; it < 300; ++it) {
        uint64_t i = rnd(&st, n), j = rnd(&st, n);
        double dx = p[2 * j] - p[2 * i], dy = p[2 * j + 1] - p[2 * i + 1];
        double len = norm2(dx, dy);
        if (i == j || isnan(len) || len <= 0.0) continue;
        double nx = -dy / len, ny = dx / len;
        size_t c = 0;
        for (size_t q = 0; q < n; ++q) {
            double ex = p[2 * q] - p[2 * i], ey = p[2 * q + 1] - p[2 * i + 1];
            if (fabs(ex * nx + ey * ny) < th_px) ++c;
        }
        if (c > best) best = c;
    }
    return n - best;
}

/* two_view.rs:605-642: DLT on (already normalised) points. */
static int homography_dlt_core(const double* a, const double* b, size_t n, double* H) {
    if (n < 4) return 0;
    double ata[81];
    memset(ata, 0, sizeof ata);
    for (size_t i = 0; i < n; ++i) {
        double x = a[2 * i], y = a[2 * i + 1], u = b[2 * i], v = b[2 * i + 1];
        double r1[9] = {-x, -y, -1.0, 0.0, 0.0, 0.0, u * x, u * y, u};
        double r2[9] = {0.0, 0.0, 0.0, -x, -y, -1.0, v * x, v * y, v};
        for (int r = 0; r < 9; ++r)
            for (int c = 0; c < 9; ++c)
                ata[9 * r + c] = ata[9 * r + c] + (r1[r] * r1[c] + r2[r] * r2[c]);
    }
    double ev[9], vec[9];
    eigen_min9(ata, ev, vec);
    for (int i = 0; i < 9; ++i) {
        H[i] = vec[i];
        if (!isfinite(H[i])) return 0;
    }
    return 1;
}

/* The homography of four correspondences of normalised points (a_i -> b_i,
 * i < 4), the exact solution of their DLT system (8 equations for H's 8
 * degrees of freedom: the null vector that homography_dlt_core finds as the
 * smallest eigenvector of A^T A, two_view.rs:607-642), in closed form
 * (Hartley & Zisserman, Multiple View Geometry, 2nd ed., ex. 4.1 / the
 * projective basis construction): with p_i = (a_i, 1), q_i = (b_i, 1),
 * M = [p0 p1 p2] and N = [q0 q1 q2], adj(M) has the rows p1 x p2, p2 x p0,
 * p0 x p1; l = adj(M) p3 and m = adj(N) q3 are the basis weights up to
 * det(M) and det(N), and H = N diag(m_i / l_i) adj(M) up to scale
 * (H p_j ~ q_j for j < 3 and H p3 ~ N m ~ q3). H[r][c] = sum_i (q_i[r] w_i)
 * adj_i[c] in index order, w_i = m_i / l_i. Returns 0 (no hypothesis) if
 * some l_i is 0 or H is not finite: three of the four points collinear. */
static void cross3(const double* u, const double* v, double* w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}
static int homography_4pt_core(const double* a, const double* b, double* H) {
    double p[4][3], q[4][3], am[3][3], an[3][3], l[3], m[3], w[3];
    for (int i = 0; i < 4; ++i) {
        p[i][0] = a[2 * i]; p[i][1] = a[2 * i + 1]; p[i][2] = 1.0;
        q[i][0] = b[2 * i]; q[i][1] = b[2 * i + 1]; q[i][2] = 1.0;
    }
    cross3(p[1], p[2], am[0]); cross3(p[2], p[0], am[1]); cross3(p[0], p[1], am[2]);
    cross3(q[1], q[2], an[0]); cross3(q[2], q[0], an[1]); cross3(q[0], q[1], an[2]);
    for (int i = 0; i < 3; ++i) {
        l[i] = am[i][0] * p[3][0] + am[i][1] * p[3][1] + am[i][2] * p[3][2];
        m[i] = an[i][0] * q[3][0] + an[i][1] * q[3][1] + an[i][2] * q[3][2];
        if (l[i] == 0.0) return 0;
        w[i] = m[i] / l[i];
    }
    for (int r = 0; r < 3; ++r) {
        const double t0 = q[0][r] * w[0], t1 = q[1][r] * w[1], t2 = q[2][r] * w[2];
        for (int c = 0; c < 3; ++c) {
            H[3 * r + c] = t0 * am[0][c] + t1 * am[1][c] + t2 * am[2][c];
            if (!isfinite(H[3 * r + c])) return 0;
        }
    }
    return 1;
}

/* matching.rs:251-266: Hartley-normalised pixel DLT. */
static int homography_dlt_x(const double* x1, const double* x2, size_t n, double* H, int four) {
    if (n < 4) return 0;
    double t1[9], t2[9];
    orc_normalizer(x1, n, t1);
    orc_normalizer(x2, n, t2);
    double* a = (double*)malloc(sizeof(double) * 2 * n);
    double* b = (double*)malloc(sizeof(double) * 2 * n);
    for (size_t i = 0; i < n; ++i) {
        double p[3] = {x1[2 * i], x1[2 * i + 1], 1.0}, v[3];
        mv3(t1, p, v);
        a[2 * i] = v[0] / v[2]; a[2 * i + 1] = v[1] / v[2];
        double q[3] = {x2[2 * i], x2[2 * i + 1], 1.0};
        mv3(t2, q, v);
        b[2 * i] = v[0] / v[2]; b[2 * i + 1] = v[1] / v[2];
    }
    double hn[9];
    int ok = four ? homography_4pt_core(a, b, hn) : homography_dlt_core(a, b, n, hn);
    free(a);
    free(b);
    if (!ok) return 0;
    double t2i[9], tmp[9], h[9];
    if (!inv3(t2, t2i)) return 0;
    mm3(t2i, hn, tmp);
    mm3(tmp, t1, h);
    double nn = fro3(h);
    if (!(isfinite(nn) && nn > 0.0)) return 0;
    for (int i = 0; i < 9; ++i) H[i] = h[i] / nn;
    return 1;
}

static int homography_dlt(const double* x1, const double* x2, size_t n, double* H) {
    return homography_dlt_x(x1, x2, n, H, 0);
}

/* The sampled hypotheses' homography (matching.rs:334-337): homography_dlt of
 * the four samples, its DLT system solved in closed form. */
static int homography_4pt(const double* x1, const double* x2, double* H) {
    return homography_dlt_x(x1, x2, 4, H, 1);
}

int orc_homography_dlt(const double* x1, const double* x2, size_t n, double* H) {
    return homography_dlt(x1, x2, n, H);
}
int orc_homography_4pt(const double* x1, const double* x2, double* H) {
    return homography_4pt(x1, x2, H);
}

/* matching.rs:269-287. */
static int htr(const double* m, const double* x, double* out) {
    double p[3] = {x[0], x[1], 1.0}, v[3];
    mv3(m, p, v);
    if (fabs(v[2]) < 1e-12) return 0;
    out[0] = v[0] / v[2];
    out[1] = v[1] / v[2];
    return 1;
}

static double homography_error(const double* h, const double* hi, const double* p,
                               const double* q) {
    double a[2], b[2];
    if (!htr(h, p, a) || !htr(hi, q, b)) return INFINITY;
    double e1 = norm2(a[0] - q[0], a[1] - q[1]);
    double e2 = norm2(b[0] - p[0], b[1] - p[1]);
    return fmax(e1, e2);
}

static size_t count_h(const double* h, const double* hi, const double* x1, const double* x2,
                      size_t n, double th) {
    size_t c = 0;
    for (size_t i = 0; i < n; ++i)
        if (homography_error(h, hi, x1 + 2 * i, x2 + 2 * i) < th) ++c;
    return c;
}

/* matching.rs:300-369. Returns 1 with H, H^-1 and the count. */
static int fit_homography(const double* x1, const double* x2, size_t n, double th_px, double* H,
                          double* Hi, size_t* cnt_out) {
    if (n < 4) return 0;
    uint64_t st = 0x2545F4914F6CDD1DULL;
    int have = 0;
    double bh[9], bhi[9];
    size_t bc = 0;
    uint64_t iters = 500, it = 0;
    double s1[8], s2[8];
    while (it < iters) {
        it++;
        uint64_t idx[4];
        int k = 0;
        while (k < 4) {
            uint64_t c = rnd(&st, n);
            int dup = 0;
            for (int e = 0; e < k; ++e) dup |= idx[e] == c;
            if (!dup) idx[k++] = c;
        }
        for (int e = 0; e < 4; ++e) {
            s1[2 * e] = x1[2 * idx[e]]; s1[2 * e + 1] = x1[2 * idx[e] + 1];
            s2[2 * e] = x2[2 * idx[e]]; s2[2 * e + 1] = x2[2 * idx[e] + 1];
        }
        double h[9], hi[9];
        if (!homography_4pt(s1, s2, h)) continue;
        if (!inv3(h, hi)) continue;
        size_t c = count_h(h, hi, x1, x2, n, th_px);
        if (!have || c > bc) {
            have = 1;
            memcpy(bh, h, sizeof bh);
            memcpy(bhi, hi, sizeof bhi);
            bc = c;
            iters = orc_adaptive_iterations((double)c / (double)n, 4, 0.999, 500);
        }
    }
    if (!have) return 0;
    double ch[9], chi[9];
    memcpy(ch, bh, sizeof ch);
    memcpy(chi, bhi, sizeof chi);
    static const double HM[4] = {4.0, 2.0, 1.0, 1.0};
    double* sx1 = (double*)malloc(sizeof(double) * 2 * n);
    double* sx2 = (double*)malloc(sizeof(double) * 2 * n);
    for (int r = 0; r < 4; ++r) {
        size_t ns = 0;
        for (size_t i = 0; i < n; ++i) {
            if (homography_error(ch, chi, x1 + 2 * i, x2 + 2 * i) < th_px * HM[r]) {
                sx1[2 * ns] = x1[2 * i]; sx1[2 * ns + 1] = x1[2 * i + 1];
                sx2[2 * ns] = x2[2 * i]; sx2[2 * ns + 1] = x2[2 * i + 1];
                ns++;
            }
        }
        double g[9], gi[9];
        if (!homography_dlt(sx1, sx2, ns, g)) break;
        if (!inv3(g, gi)) break;
        memcpy(ch, g, sizeof ch);
        memcpy(chi, gi, sizeof chi);
        size_t c = count_h(g, gi, x1, x2, n, th_px);
        if (c > bc) {
            memcpy(bh, g, sizeof bh);
            memcpy(bhi, gi, sizeof bhi);
            bc = c;
        }
    }
    free(sx1);
    free(sx2);
    memcpy(H, bh, sizeof bh);
    memcpy(Hi, bhi, sizeof bhi);
    *cnt_out = bc;
    return 1;
}

uint64_t orc_homography_support(const double* x1, const double* x2, size_t n, double th_px) {
    double h[9], hi[9];
    size_t c = 0;
    if (!fit_homography(x1, x2, n, th_px, h, hi, &c)) return 0;
    return c;
}

/* matching.rs:405-409. */
uint64_t orc_parallax_needed(uint64_t n) {
    for (uint64_t k = 2; k <= n; ++k)
        if (orc_binomial_tail(n, k, 0.001) < 1e-6) return k;
    return n + 1;
}

/* matching.rs:425-492. */
int orc_select_two_view_model(const double* x1, const double* x2, size_t n, const double* F,
                              double sigma_px, orc_model_sel* out) {
    memset(out, 0, sizeof *out);
    if (n < 8 || !isfinite(sigma_px) || sigma_px <= 0.0 || !all_finite2(x1, n) ||
        !all_finite2(x2, n))
        return 0;
    for (int i = 0; i < 9; ++i)
        if (!isfinite(F[i])) return 0;
    double* es = (double*)malloc(sizeof(double) * n);
    size_t ne = 0;
    for (size_t i = 0; i < n; ++i) {
        double e = orc_sampson_error(F, x1 + 2 * i, x2 + 2 * i);
        if (isfinite(e)) es[ne++] = e;
    }
    if (ne < 8) { free(es); return 0; }
    qsort(es, ne, sizeof(double), cmp_double);
    double sigma_hat = sqrt(es[ne / 2] / 0.455);
    free(es);
    double sigma = fmax(sigma_px, sigma_hat);
    double s2 = sigma * sigma;
    double h[9], hi[9];
    size_t hc;
    if (!fit_homography(x1, x2, n, 3.0 * sigma, h, hi, &hc)) return 0;
    double rho_f = 0.0, rho_h = 0.0;
    uint64_t parallax = 0;
    for (size_t i = 0; i < n; ++i) {
        const double* p = x1 + 2 * i;
        const double* q = x2 + 2 * i;
        double ef = orc_sampson_error(F, p, q) / s2;
        double a[2], b[2], t1, t2;
        t1 = htr(h, p, a) ? (a[0] - q[0]) * (a[0] - q[0]) + (a[1] - q[1]) * (a[1] - q[1]) : INFINITY;
        t2 = htr(hi, q, b) ? (b[0] - p[0]) * (b[0] - p[0]) + (b[1] - p[1]) * (b[1] - p[1]) : INFINITY;
        double eh = 0.25 * (t1 + t2) / s2;
        rho_f += fmin(ef, 2.0 * (4.0 - 3.0));
        rho_h += fmin(eh, 2.0 * (4.0 - 2.0));
        if (eh > 13.8 && ef < 10.8) parallax++;
    }
    double nf = (double)n;
    double l1 = log(4.0), l2 = log(4.0 * nf);
    out->gric_f = rho_f + l1 * 3.0 * nf + l2 * 7.0;
    out->gric_h = rho_h + l1 * 2.0 * nf + l2 * 8.0;
    out->parallax = parallax;
    out->parallax_needed = orc_parallax_needed(n);
    out->model = (out->gric_h < out->gric_f && parallax < out->parallax_needed) ? 1 : 0;
    out->sigma_px = sigma;
    out->has_selection = 1;
    return 1;
}

void orc_ransac_cfg_default(orc_ransac_cfg* c) {
    c->threshold_px = 1.5;
    c->min_inlier_ratio = 0.2;
    c->max_iters = 2000;
    c->confidence = 0.999;
    c->seed = 1;
    c->min_inliers = 15;
    c->max_chance_prob = 1e-6;
}

static size_t inliers_of(const double* F, const double* x1, const double* x2, size_t n,
                         double th2, uint8_t* mask) {
    size_t c = 0;
    for (size_t i = 0; i < n; ++i) {
        mask[i] = orc_sampson_error(F, x1 + 2 * i, x2 + 2 * i) < th2;
        c += mask[i];
    }
    return c;
}

static size_t select_pts(const double* F, const double* x1, const double* x2, size_t n,
                         double t2, double* s1, double* s2) {
    size_t k = 0;
    for (size_t i = 0; i < n; ++i) {
        if (orc_sampson_error(F, x1 + 2 * i, x2 + 2 * i) < t2) {
            s1[2 * k] = x1[2 * i]; s1[2 * k + 1] = x1[2 * i + 1];
            s2[2 * k] = x2[2 * i]; s2[2 * k + 1] = x2[2 * i + 1];
            k++;
        }
    }
    return k;
}

static size_t gather_mask(const uint8_t* mask, const double* x1, const double* x2, size_t n,
                          double* s1, double* s2) {
    size_t k = 0;
    for (size_t i = 0; i < n; ++i) {
        if (mask[i]) {
            s1[2 * k] = x1[2 * i]; s1[2 * k + 1] = x1[2 * i + 1];
            s2[2 * k] = x2[2 * i]; s2[2 * k + 1] = x2[2 * i + 1];
            k++;
        }
    }
    return k;
}

typedef struct {
    double F[9];
    uint8_t* mask;
    size_t cnt;
    uint64_t origin; /* iteration whose local optimisation produced it (trace only) */
} hyp_t;

/* Hypothesis-loop trace (tests only; see oracle.h). */
static __thread orc_trace_rec* g_tr = NULL;
static __thread size_t g_tr_cap = 0, g_tr_len = 0, g_tr_n = 0;
static __thread uint8_t* g_tr_masks = NULL;

void orc_ransac_trace_begin(orc_trace_rec* recs, size_t cap, uint8_t* lo_masks, size_t n) {
    g_tr = recs;
    g_tr_cap = cap;
    g_tr_len = 0;
    g_tr_n = n;
    g_tr_masks = lo_masks;
}

size_t orc_ransac_trace_end(void) {
    g_tr = NULL;
    return g_tr_len;
}

/* Closes the current record: the limit and pool after the iteration. */
static void trace_close(orc_trace_rec* r, uint64_t iters, const hyp_t* pool, int npool) {
    if (!r) return;
    r->iters_after = iters;
    r->npool = npool;
    for (int k = 0; k < npool && k < 5; ++k) {
        r->pool_cnt[k] = pool[k].cnt;
        r->pool_origin[k] = pool[k].origin;
    }
    g_tr_len++;
}

void orc_lm_work(uint64_t* iters, uint64_t* evals) {
    *iters = g_lm_iters;
    *evals = g_lm_evals;
}

void orc_ransac_last_stats(uint64_t* iters, uint64_t* lo_runs, uint64_t* lm_calls,
                           uint64_t* lm_converged, uint64_t* lm_stalled_at_start) {
    if (iters) *iters = g_iters;
    if (lo_runs) *lo_runs = g_lo_runs;
    if (lm_calls) *lm_calls = g_lm_calls;
    if (lm_converged) *lm_converged = g_lm_conv;
    if (lm_stalled_at_start) *lm_stalled_at_start = g_lm_stall0;
}

/* matching.rs:941-1104. */
int orc_ransac_fundamental(const double* x1, const double* x2, size_t n, const orc_ransac_cfg* cfg,
                           double* F_out, uint8_t* inl_out, orc_model_sel* sel) {
    g_iters = g_lo_runs = g_lm_calls = g_lm_conv = g_lm_stall0 = 0;
    g_lm_iters = g_lm_evals = 0;
    if (sel) memset(sel, 0, sizeof *sel);
    if (n < 8 || !all_finite2(x1, n) || !all_finite2(x2, n)) return 0;
    const double th2 = cfg->threshold_px * cfg->threshold_px;
    uint64_t st = cfg->seed ^ 0x9E3779B97F4A7C15ULL;
    double* s1 = (double*)malloc(sizeof(double) * 2 * n);
    double* s2 = (double*)malloc(sizeof(double) * 2 * n);
    uint8_t* inl = (uint8_t*)malloc(n);
    uint8_t* gi = (uint8_t*)malloc(n);
    uint8_t* best_mask = (uint8_t*)malloc(n);
    hyp_t pool[6];
    for (int k = 0; k < 6; ++k) pool[k].mask = (uint8_t*)malloc(n);
    int npool = 0;
    int have_best = 0;
    double bestF[9];
    size_t best_cnt_v = 0;
    uint64_t iters = cfg->max_iters, it = 0;
    while (it < iters) {
        it++;
        orc_trace_rec* tr = (g_tr && g_tr_len < g_tr_cap && g_tr_n == n) ? g_tr + g_tr_len : NULL;
        if (tr) memset(tr, 0, sizeof *tr);
        uint64_t idx[8];
        int k = 0;
        while (k < 8) {
            uint64_t c = rnd(&st, n);
            int dup = 0;
            for (int e = 0; e < k; ++e) dup |= idx[e] == c;
            if (!dup) idx[k++] = c;
        }
        double a8[16], b8[16];
        for (int e = 0; e < 8; ++e) {
            a8[2 * e] = x1[2 * idx[e]]; a8[2 * e + 1] = x1[2 * idx[e] + 1];
            b8[2 * e] = x2[2 * idx[e]]; b8[2 * e + 1] = x2[2 * idx[e] + 1];
        }
        size_t best_cnt = have_best ? best_cnt_v : 0;
        if (tr) {
            tr->it = it;
            memcpy(tr->idx, idx, sizeof idx);
            tr->best_before = best_cnt;
        }
        double f[9];
        if (!orc_fundamental_8pt(a8, b8, 8, f)) {
            trace_close(tr, iters, pool, npool);
            continue;
        }
        size_t cnt = inliers_of(f, x1, x2, n, th2, inl);
        if (tr) {
            tr->ok = 1;
            memcpy(tr->F, f, sizeof f);
            tr->raw_cnt = tr->lo_cnt = cnt;
        }
        if (cnt >= 8 && 4 * cnt >= best_cnt) {
            g_lo_runs++;
            double f_lo[9], cur[9];
            memcpy(f_lo, f, sizeof f_lo);
            memcpy(cur, f, sizeof cur);
            for (int r = 0; r < 6; ++r) {
                double t2 = th2 * NM[r] * NM[r];
                size_t ns = select_pts(cur, x1, x2, n, t2, s1, s2);
                double g[9];
                if (tr) tr->stage_ns[r] = ns;
                if (!orc_fundamental_8pt(s1, s2, ns, g)) break;
                memcpy(cur, g, sizeof cur);
                size_t gc = inliers_of(g, x1, x2, n, th2, gi);
                if (tr) {
                    tr->stage_ok[r] = 1;
                    tr->stage_cnt[r] = gc;
                    memcpy(tr->stage_F[r], g, sizeof g);
                }
                if (gc > cnt) {
                    memcpy(f_lo, g, sizeof f_lo);
                    memcpy(inl, gi, n);
                    cnt = gc;
                }
            }
            if (tr) {
                tr->lo = 1;
                tr->lo_cnt = cnt;
                memcpy(tr->F_lo, f_lo, sizeof f_lo);
                if (g_tr_masks) memcpy(g_tr_masks + g_tr_len * n, inl, n);
            }
            /* matching.rs:1017-1028: Jaccard > 0.7 valley test, stable sort. */
            int same = -1;
            for (int j = 0; j < npool && same < 0; ++j) {
                size_t both = 0, either = 0;
                for (size_t i = 0; i < n; ++i) {
                    both += pool[j].mask[i] && inl[i];
                    either += pool[j].mask[i] || inl[i];
                }
                if (10 * both > 7 * either) same = j;
            }
            if (same >= 0) {
                if (pool[same].cnt < cnt) {
                    memcpy(pool[same].F, f_lo, sizeof f_lo);
                    memcpy(pool[same].mask, inl, n);
                    pool[same].cnt = cnt;
                    pool[same].origin = it;
                }
            } else {
                memcpy(pool[npool].F, f_lo, sizeof f_lo);
                memcpy(pool[npool].mask, inl, n);
                pool[npool].cnt = cnt;
                pool[npool].origin = it;
                npool++;
            }
            /* stable insertion sort by count, descending */
            for (int a = 1; a < npool; ++a) {
                hyp_t x = pool[a];
                int b = a - 1;
                while (b >= 0 && pool[b].cnt < x.cnt) {
                    pool[b + 1] = pool[b];
                    b--;
                }
                pool[b + 1] = x;
            }
            if (npool > 5) npool = 5; /* the dropped mask buffer stays at pool[5] */
            if (cnt <= best_cnt) {
                trace_close(tr, iters, pool, npool);
                continue;
            }
            double w = (double)cnt / (double)n;
            iters = orc_adaptive_iterations(w, 8, cfg->confidence, cfg->max_iters);
            memcpy(bestF, f_lo, sizeof bestF);
            memcpy(best_mask, inl, n);
            best_cnt_v = cnt;
            have_best = 1;
        }
        trace_close(tr, iters, pool, npool);
    }
    g_iters = it;
    int result = 0;
    if (!have_best) goto done;
    {
        double f[9];
        memcpy(f, bestF, sizeof f);
        memcpy(inl, best_mask, n);
        size_t inl_cnt = best_cnt_v;
        /* matching.rs:1041-1057 */
        for (int p = 0; p < npool; ++p) {
            double cur[9];
            memcpy(cur, pool[p].F, sizeof cur);
            for (int r = 0; r < 6; ++r) {
                double t2 = th2 * NM[r] * NM[r];
                size_t ns = select_pts(cur, x1, x2, n, t2, s1, s2);
                double nxt[9];
                orc_refine_sampson(cur, s1, s2, ns, 30, nxt, NULL, NULL);
                memcpy(cur, nxt, sizeof cur);
                size_t c = inliers_of(cur, x1, x2, n, th2, gi);
                if (c > inl_cnt) {
                    memcpy(f, cur, sizeof f);
                    memcpy(inl, gi, n);
                    inl_cnt = c;
                }
            }
        }
        /* matching.rs:1059-1069 */
        for (int r = 0; r < 2; ++r) {
            size_t ns = gather_mask(inl, x1, x2, n, s1, s2);
            double g[9];
            orc_refine_sampson(f, s1, s2, ns, 30, g, NULL, NULL);
            size_t c = inliers_of(g, x1, x2, n, th2, gi);
            if (c < ns) break;
            memcpy(f, g, sizeof f);
            memcpy(inl, gi, n);
            inl_cnt = c;
        }
        size_t cnt = inl_cnt;
        /* matching.rs:1071-1091 */
        size_t minc = cfg->min_inliers > 8 ? cfg->min_inliers : 8;
        if (cnt < minc || (double)cnt < cfg->min_inlier_ratio * (double)n) goto done;
        if (cfg->max_chance_prob > 0.0) {
            double p = orc_epipolar_band_probability(x2, n, cfg->threshold_px);
            if (orc_chance_inlier_probability(n, cnt, p) > cfg->max_chance_prob) goto done;
        }
        size_t ns = gather_mask(inl, x1, x2, n, s1, s2);
        if (nearly_collinear(s1, ns) || nearly_collinear(s2, ns)) goto done;
        if (off_line_count(s1, ns, cfg->threshold_px) < 8 ||
            off_line_count(s2, ns, cfg->threshold_px) < 8)
            goto done;
        /* matching.rs:1096-1097 */
        orc_model_sel ms;
        int has = orc_select_two_view_model(s1, s2, ns, f, cfg->threshold_px / 3.0, &ms);
        if (sel) {
            if (has) *sel = ms;
            else memset(sel, 0, sizeof *sel);
        }
        memcpy(F_out, f, sizeof f);
        memcpy(inl_out, inl, n);
        result = 1;
    }
done:
    free(s1);
    free(s2);
    free(inl);
    free(gi);
    free(best_mask);
    for (int k = 0; k < 6; ++k) free(pool[k].mask);
    return result;
}
