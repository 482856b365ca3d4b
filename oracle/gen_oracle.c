/*
 * Test-geometry generators of the reference's RANSAC tests. TEST
 * INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates crates/core/src/matching.rs:1165-1220 (two_cameras,
 * correspondences), :1792-1806 (noisy_lines), :1825-1845
 * (planar_correspondences), :1925-1952 (ground_with_buildings), :2201-2230
 * (line_plus_general), :1107-1133 (fundamental_from_cameras), with the
 * pinhole pieces of crates/core/src/camera.rs:41-52 (from_hfov), :123-159
 * (Pose::from_center, project) and math.rs:8-10 (skew). nalgebra's
 * Rotation3::from_euler_angles(roll, pitch, yaw) = Rz(yaw) Ry(pitch) Rx(roll).
 *
 * Camera layout (18 doubles): R row-major (9), t (3), fx, fy, cx, cy, w, h.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

static void euler(double roll, double pitch, double yaw, double* R) {
    double sr = sin(roll), cr = cos(roll);
    double sp = sin(pitch), cp = cos(pitch);
    double sy = sin(yaw), cy = cos(yaw);
    double m[9] = {cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr,
                   sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr,
                   -sp,     cp * sr,                cp * cr};
    memcpy(R, m, sizeof m);
}

static void mv3g(const double* m, const double* x, double* y) {
    double a = m[0] * x[0] + m[1] * x[1] + m[2] * x[2];
    double b = m[3] * x[0] + m[4] * x[1] + m[5] * x[2];
    double c = m[6] * x[0] + m[7] * x[1] + m[8] * x[2];
    y[0] = a; y[1] = b; y[2] = c;
}

static void mm3g(const double* A, const double* B, double* C) {
    double t[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            t[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
    memcpy(C, t, sizeof t);
}

static void make_cam(double roll, double pitch, double yaw, const double* centre, double* cam) {
    double R[9];
    euler(roll, pitch, yaw, R);
    double rc[3];
    mv3g(R, centre, rc);
    memcpy(cam, R, sizeof R);
    cam[9] = -rc[0];
    cam[10] = -rc[1];
    cam[11] = -rc[2];
    /* from_hfov(960, 540, 70 deg) (camera.rs:41-52), to_radians = x * (pi / 180) */
    double hfov = 70.0 * (3.141592653589793 / 180.0);
    double f = 0.5 * 960.0 / tan(0.5 * hfov);
    cam[12] = f;
    cam[13] = f;
    cam[14] = 0.5 * 960.0;
    cam[15] = 0.5 * 540.0;
    cam[16] = 960.0;
    cam[17] = 540.0;
}

void orc_two_cameras(double* c1, double* c2) {
    const double C1[3] = {0.0, 0.0, -10.0}, C2[3] = {2.0, 0.5, -10.3};
    make_cam(0.05, -0.02, 0.1, C1, c1);
    make_cam(0.12, 0.08, -0.05, C2, c2);
}

/* camera.rs:146-155 (zero distortion: fx * x/z + cx). */
int orc_project(const double* cam, const double* X, double* px) {
    double xc[3];
    mv3g(cam, X, xc);
    xc[0] = xc[0] + cam[9];
    xc[1] = xc[1] + cam[10];
    xc[2] = xc[2] + cam[11];
    if (xc[2] <= 0.0) return 0;
    px[0] = cam[12] * (xc[0] / xc[2]) + cam[14];
    px[1] = cam[13] * (xc[1] / xc[2]) + cam[15];
    return 1;
}

static int contains(const double* cam, const double* p) {
    return p[0] >= 0.0 && p[1] >= 0.0 && p[0] < cam[16] && p[1] < cam[17];
}

/* matching.rs:1107-1133. */
void orc_fundamental_from_cameras(const double* c1, const double* c2, double* F) {
    double R1t[9] = {c1[0], c1[3], c1[6], c1[1], c1[4], c1[7], c1[2], c1[5], c1[8]};
    double r[9];
    mm3g(c2, R1t, r);
    double rt1[3];
    mv3g(r, c1 + 9, rt1);
    double t[3] = {c2[9] - rt1[0], c2[10] - rt1[1], c2[11] - rt1[2]};
    double sk[9] = {0.0, -t[2], t[1], t[2], 0.0, -t[0], -t[1], t[0], 0.0};
    double e[9];
    mm3g(sk, r, e);
    double k1[9], k2[9];
    const double* cs[2] = {c1, c2};
    double* ks[2] = {k1, k2};
    for (int q = 0; q < 2; ++q) {
        const double* c = cs[q];
        double ox = (0.0 - c[14]) / c[12], oy = (0.0 - c[15]) / c[13];
        double exx = (1.0 - c[14]) / c[12], exy = (0.0 - c[15]) / c[13];
        double eyx = (0.0 - c[14]) / c[12], eyy = (1.0 - c[15]) / c[13];
        double k[9] = {exx - ox, eyx - ox, ox, exy - oy, eyy - oy, oy, 0.0, 0.0, 1.0};
        memcpy(ks[q], k, sizeof k);
    }
    double k2t[9] = {k2[0], k2[3], k2[6], k2[1], k2[4], k2[7], k2[2], k2[5], k2[8]};
    double tmp[9], f[9];
    mm3g(k2t, e, tmp);
    mm3g(tmp, k1, f);
    double s = 0.0;
    for (int c = 0; c < 3; ++c) s = s + (f[c] * f[c] + f[3 + c] * f[3 + c] + f[6 + c] * f[6 + c]);
    double n = sqrt(s);
    for (int i = 0; i < 9; ++i) F[i] = f[i] / n;
}

/* matching.rs:1182-1220. */
void orc_correspondences(size_t n, double sigma, double out, uint64_t seed, double* x1, double* x2,
                         uint8_t* truth) {
    double c1[18], c2[18];
    orc_two_cameras(c1, c2);
    orc_lcg g = {seed};
    size_t k = 0;
    while (k < n) {
        double X[3];
        X[0] = orc_lcg_next(&g) * 12.0 - 6.0;
        X[1] = orc_lcg_next(&g) * 8.0 - 4.0;
        X[2] = orc_lcg_next(&g) * 6.0 - 3.0;
        double a[2], b[2];
        if (!orc_project(c1, X, a) || !orc_project(c2, X, b)) continue;
        if (!contains(c1, a) || !contains(c2, b)) continue;
        int outlier = orc_lcg_next(&g) < out;
        if (outlier) {
            b[0] = orc_lcg_next(&g) * 960.0;
            b[1] = orc_lcg_next(&g) * 540.0;
        } else {
            double gx = orc_lcg_gauss(&g), gy = orc_lcg_gauss(&g);
            b[0] = b[0] + gx * sigma;
            b[1] = b[1] + gy * sigma;
        }
        double gx = orc_lcg_gauss(&g), gy = orc_lcg_gauss(&g);
        x1[2 * k] = a[0] + gx * sigma;
        x1[2 * k + 1] = a[1] + gy * sigma;
        x2[2 * k] = b[0];
        x2[2 * k + 1] = b[1];
        if (truth) truth[k] = !outlier;
        ++k;
    }
}

static void push_noisy(orc_lcg* g, const double* a, const double* b, double sigma, double* x1,
                       double* x2) {
    double gx = orc_lcg_gauss(g), gy = orc_lcg_gauss(g);
    x1[0] = a[0] + gx * sigma;
    x1[1] = a[1] + gy * sigma;
    gx = orc_lcg_gauss(g);
    gy = orc_lcg_gauss(g);
    x2[0] = b[0] + gx * sigma;
    x2[1] = b[1] + gy * sigma;
}

/* matching.rs:1825-1845. */
void orc_planar_correspondences(size_t n, double sigma, uint64_t seed, double* x1, double* x2) {
    double c1[18], c2[18];
    orc_two_cameras(c1, c2);
    orc_lcg g = {seed};
    size_t k = 0;
    while (k < n) {
        double X[3];
        X[0] = orc_lcg_next(&g) * 14.0 - 7.0;
        X[1] = orc_lcg_next(&g) * 8.0 - 4.0;
        X[2] = 0.0;
        double a[2], b[2];
        if (!orc_project(c1, X, a) || !orc_project(c2, X, b)) continue;
        if (!contains(c1, a) || !contains(c2, b)) continue;
        push_noisy(&g, a, b, sigma, x1 + 2 * k, x2 + 2 * k);
        ++k;
    }
}

/* matching.rs:1925-1952. */
void orc_ground_with_buildings(size_t n, double bld, double sigma, uint64_t seed, double* x1,
                               double* x2) {
    double c1[18], c2[18];
    orc_two_cameras(c1, c2);
    orc_lcg g = {seed};
    size_t nb = (size_t)round((double)n * bld);
    size_t k = 0;
    while (k < n) {
        double z = k < nb ? -1.0 - 2.0 * orc_lcg_next(&g) : 0.0;
        double X[3];
        X[0] = orc_lcg_next(&g) * 14.0 - 7.0;
        X[1] = orc_lcg_next(&g) * 8.0 - 4.0;
        X[2] = z;
        double a[2], b[2];
        if (!orc_project(c1, X, a) || !orc_project(c2, X, b)) continue;
        if (!contains(c1, a) || !contains(c2, b)) continue;
        push_noisy(&g, a, b, sigma, x1 + 2 * k, x2 + 2 * k);
        ++k;
    }
}

/* matching.rs:1792-1806. */
void orc_noisy_lines(double sigma, uint64_t seed, double* l1, double* l2) {
    orc_lcg g = {seed};
    const double A[2] = {200.0, 350.0}, B[2] = {3.0, -2.0};
    double* out[2] = {l1, l2};
    for (int q = 0; q < 2; ++q) {
        for (int i = 0; i < 60; ++i) {
            double t = (double)i + orc_lcg_next(&g);
            double x = 100.0 + 13.0 * t, y = A[q] + B[q] * t;
            double gx = orc_lcg_gauss(&g), gy = orc_lcg_gauss(&g);
            out[q][2 * i] = x + gx * sigma;
            out[q][2 * i + 1] = y + gy * sigma;
        }
    }
}

/* matching.rs:2201-2230. */
void orc_line_plus_general(size_t k, double sigma, uint64_t seed, double* x1, double* x2,
                           double* h1, double* h2) {
    double c1[18], c2[18];
    orc_two_cameras(c1, c2);
    orc_lcg g = {seed};
    size_t m = 0;
    while (m < 60) {
        double t = orc_lcg_next(&g) * 12.0 - 6.0;
        double X[3] = {t, 0.3 * t + 0.5, 0.1 * t};
        double a[2], b[2];
        if (!orc_project(c1, X, a) || !orc_project(c2, X, b)) continue;
        if (!contains(c1, a) || !contains(c2, b)) continue;
        push_noisy(&g, a, b, sigma, x1 + 2 * m, x2 + 2 * m);
        ++m;
    }
    size_t tot = k + 100;
    double g1[2 * 1200], g2[2 * 1200];
    if (tot > 1200) return;
    orc_correspondences(tot, sigma, 0.0, seed + 1000, g1, g2, NULL);
    memcpy(x1 + 120, g1, sizeof(double) * 2 * k);
    memcpy(x2 + 120, g2, sizeof(double) * 2 * k);
    memcpy(h1, g1 + 2 * k, sizeof(double) * 200);
    memcpy(h2, g2 + 2 * k, sizeof(double) * 200);
}
