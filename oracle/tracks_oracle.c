/*
 * Oracle (test infrastructure only) for the multi-view track builder of
 * skylens-stream: crates/core/src/tracks.rs:127-800 (build_tracks, the
 * displacement filter, the union-find with image-conflict joins,
 * select_tracks). Only tests/ load it, as the checker of the HIP/C++ path.
 *
 * Every function follows the cited lines statement by statement. Two
 * choices of the reference are not restatable bit for bit and are pinned
 * here instead:
 *   - the k = 24 nearest candidates of the displacement filter come from
 *     Rust's select_nth_unstable_by (tracks.rs:612-615), whose output order
 *     (and, for equal distances at the k-th place, whose set) the standard
 *     library leaves unspecified; here the candidates are ordered by
 *     (squared distance, index) and the first 24 kept. The fit's sums then
 *     run in that order, so a flag can differ from the reference only when a
 *     point's offset lies within rounding of its threshold. A test knob
 *     (orc_set_displacement_order) reorders the 24 so that the tests can
 *     check that the flags do not depend on the order.
 *   - nalgebra 0.33.3 (Cargo.lock) supplies the 3x3 determinant and inverse
 *     of fit_affine; they are restated from its published cofactor formulas
 *     (determinant: first-row expansion with the minors m22 m33 - m32 m23,
 *     m21 m33 - m31 m23, m21 m32 - m31 m22; inverse: adjugate / determinant).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define DISP_TOLERANCE 40.0   /* tracks.rs:6-11 */
#define DISP_FLOOR 0.25
#define DISP_CAP 3.0
#define DISP_REF_W 960.0
#define DISP_REF_H 540.0
#define DISP_K 24
#define VOTE_ROUNDS 3         /* tracks.rs:158 */

/* ------------------------------------------------------------ helpers */

/* f64::total_cmp on the values this file sorts (finite, signed zeros). */
static int total_cmp(double a, double b) {
    if (a < b) return -1;
    if (a > b) return 1;
    const int sa = signbit(a) != 0, sb = signbit(b) != 0;
    return sb - sa; /* -0.0 before +0.0 */
}
static int cmp_f64(const void* x, const void* y) { return total_cmp(*(const double*)x, *(const double*)y); }

/* `as i64` of a float: saturating, NaN -> 0 (Rust's float-to-int cast). */
static int64_t sat_i64(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}

/* ------------------------------------------------------ displacement filter */

typedef struct { int64_t kx, ky; uint64_t i; } cell_key;
static int cmp_cell(const void* x, const void* y) {
    const cell_key* a = x;
    const cell_key* b = y;
    if (a->kx != b->kx) return a->kx < b->kx ? -1 : 1;
    if (a->ky != b->ky) return a->ky < b->ky ? -1 : 1;
    return a->i < b->i ? -1 : (a->i > b->i);
}
typedef struct { double d2; uint64_t j; } cand_t;
static int cmp_cand(const void* x, const void* y) {
    const cand_t* a = x;
    const cand_t* b = y;
    const int c = total_cmp(a->d2, b->d2);
    if (c) return c;
    return a->j < b->j ? -1 : (a->j > b->j);
}

/* q_range of tracks.rs:552-558: the 5 % and 95 % order statistics' gap. */
static double q_range(const double* v0, size_t n, double* tmp) {
    memcpy(tmp, v0, n * sizeof(double));
    qsort(tmp, n, sizeof(double), cmp_f64);
    const size_t lo = n / 20, hi = n - 1 - n / 20;
    return fmax(tmp[hi] - tmp[lo], 0.0);
}

/* fit_affine (tracks.rs:672-689) over the kept neighbours; 0 = None. */
static int fit_affine(const double (*nb)[4], const uint8_t* keep, size_t n, double* f) {
    double m[3][3] = {{0}}, rx[3] = {0}, ry[3] = {0};
    for (size_t k = 0; k < n; ++k) {
        if (!keep[k]) continue;
        const double v[3] = {1.0, nb[k][0], nb[k][1]};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) m[r][c] = m[r][c] + v[r] * v[c];
        for (int r = 0; r < 3; ++r) rx[r] = rx[r] + v[r] * nb[k][2];
        for (int r = 0; r < 3; ++r) ry[r] = ry[r] + v[r] * nb[k][3];
    }
    const double m11 = m[0][0], m12 = m[0][1], m13 = m[0][2];
    const double m21 = m[1][0], m22 = m[1][1], m23 = m[1][2];
    const double m31 = m[2][0], m32 = m[2][1], m33 = m[2][2];
    const double minor_1 = m22 * m33 - m32 * m23;
    const double minor_2 = m21 * m33 - m31 * m23;
    const double minor_3 = m21 * m32 - m31 * m22;
    const double det = m11 * minor_1 - m12 * minor_2 + m13 * minor_3;
    if (det != det || det <= 1e-4 * (m11 * m11 * m11)) return 0;
    /* try_inverse: the same determinant, zero -> None */
    if (det == 0.0) return 0;
    double inv[3][3];
    inv[0][0] = minor_1 / det;
    inv[0][1] = (m13 * m32 - m33 * m12) / det;
    inv[0][2] = (m12 * m23 - m22 * m13) / det;
    inv[1][0] = -minor_2 / det;
    inv[1][1] = (m11 * m33 - m31 * m13) / det;
    inv[1][2] = (m13 * m21 - m23 * m11) / det;
    inv[2][0] = minor_3 / det;
    inv[2][1] = (m12 * m31 - m32 * m11) / det;
    inv[2][2] = (m11 * m22 - m21 * m12) / det;
    for (int r = 0; r < 3; ++r) {
        f[r] = inv[r][0] * rx[0] + inv[r][1] * rx[1] + inv[r][2] * rx[2];
        f[3 + r] = inv[r][0] * ry[0] + inv[r][1] * ry[1] + inv[r][2] * ry[2];
    }
    return 1;
}

/* The order in which the 24 kept candidates enter the fits (test knob): 0
 * (squared distance, index); 1 the reverse; 2 ascending index; 3 a seeded
 * shuffle. The reference leaves this order unspecified (tracks.rs:612-615),
 * so the tests check that the flags do not depend on it. */
static int g_disp_order = 0;
void orc_set_displacement_order(int mode) { g_disp_order = mode; }

static int cmp_cand_index(const void* x, const void* y) {
    const cand_t* a = x;
    const cand_t* b = y;
    return a->j < b->j ? -1 : a->j > b->j;
}

static void permute_candidates(cand_t* c, size_t nc, uint64_t seed) {
    if (g_disp_order == 1) {
        for (size_t a = 0, b = nc - 1; a < b; ++a, --b) {
            const cand_t t = c[a];
            c[a] = c[b];
            c[b] = t;
        }
    } else if (g_disp_order == 2) {
        qsort(c, nc, sizeof(cand_t), cmp_cand_index);
    } else if (g_disp_order == 3) {
        uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
        for (size_t k = nc - 1; k > 0; --k) {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            const size_t r = (size_t)(x % (k + 1));
            const cand_t t = c[k];
            c[k] = c[r];
            c[r] = t;
        }
    }
}

/* displacement_outliers_with_floor (tracks.rs:519-670). m: n (fa, fb).
 * offs (tests only, may be NULL): each match's fitted offset r and its
 * threshold, NAN where no fit ran. */
static void displacement_impl(const uint64_t* m, size_t n, const double* kpa, const double* kpb,
                              double floor_, double cap_, uint8_t* out, double* offs) {
    memset(out, 0, n);
    if (n < 8) return;
    double* px = malloc(n * sizeof(double));
    double* py = malloc(n * sizeof(double));
    double* dx = malloc(n * sizeof(double));
    double* dy = malloc(n * sizeof(double));
    double* tmp = malloc(n * sizeof(double));
    cell_key* idx = malloc(n * sizeof(cell_key));
    cand_t* cand = malloc(n * sizeof(cand_t));
    if (!px || !py || !dx || !dy || !tmp || !idx || !cand) goto done;
    double max_x = 0.0, max_y = 0.0, min_x = DBL_MAX, min_y = DBL_MAX;
    for (size_t k = 0; k < n; ++k) {
        const uint64_t fa = m[2 * k], fb = m[2 * k + 1];
        px[k] = kpa[2 * fa];
        py[k] = kpa[2 * fa + 1];
        dx[k] = kpb[2 * fb] - kpa[2 * fa];
        dy[k] = kpb[2 * fb + 1] - kpa[2 * fa + 1];
        max_x = fmax(max_x, px[k]);
        max_y = fmax(max_y, py[k]);
        min_x = fmin(min_x, px[k]);
        min_y = fmin(min_y, py[k]);
    }
    const double size = fmax(max_x / DISP_REF_W, max_y / DISP_REF_H);
    if (size <= 0.0 || !isfinite(size)) goto done;
    const double tol = DISP_TOLERANCE * size;
    const double full = fmax((max_x - min_x) * (max_y - min_y), 1e-12);
    const double qx = q_range(px, n, tmp), qy = q_range(py, n, tmp);
    const double quant = fmax(qx * qy / 0.81, 1e-12);
    const double area = full > 16.0 * quant ? quant : full;
    const double cell = fmax(fmax(sqrt(area * 8.0 / (double)n), size * 4.0), 1e-9);
    for (size_t k = 0; k < n; ++k) {
        idx[k].kx = sat_i64(floor(px[k] / cell));
        idx[k].ky = sat_i64(floor(py[k] / cell));
        idx[k].i = k;
    }
    qsort(idx, n, sizeof(cell_key), cmp_cell);
    for (size_t i = 0; i < n; ++i) {
        const int64_t cx = sat_i64(floor(px[i] / cell)), cy = sat_i64(floor(py[i] / cell));
        size_t nc = 0;
        for (int64_t ring = 1; ring <= 3; ++ring) {
            nc = 0;
            for (int64_t gx = cx - ring; gx <= cx + ring; ++gx)
                for (int64_t gy = cy - ring; gy <= cy + ring; ++gy) {
                    /* partition_point: first entry with key >= (gx, gy) */
                    size_t lo = 0, hi = n;
                    while (lo < hi) {
                        const size_t mid = (lo + hi) / 2;
                        const int less = idx[mid].kx < gx || (idx[mid].kx == gx && idx[mid].ky < gy);
                        if (less) lo = mid + 1;
                        else hi = mid;
                    }
                    for (size_t e = lo; e < n && idx[e].kx == gx && idx[e].ky == gy; ++e) {
                        const uint64_t j = idx[e].i;
                        if (j == i) continue;
                        const double ex = px[j] - px[i], ey = py[j] - py[i];
                        cand[nc].d2 = ex * ex + ey * ey;
                        cand[nc].j = j;
                        ++nc;
                    }
                }
            if (nc >= DISP_K) break;
        }
        if (nc < 8) continue;
        if (nc > DISP_K) {
            qsort(cand, nc, sizeof(cand_t), cmp_cand);
            nc = DISP_K;
            if (g_disp_order) permute_candidates(cand, nc, i);
        }
        double nb[DISP_K][4], res[DISP_K], sorted[DISP_K], f[6] = {0};
        uint8_t keep[DISP_K];
        for (size_t k = 0; k < nc; ++k) {
            const uint64_t j = cand[k].j;
            nb[k][0] = (px[j] - px[i]) / cell;
            nb[k][1] = (py[j] - py[i]) / cell;
            nb[k][2] = dx[j] - dx[i];
            nb[k][3] = dy[j] - dy[i];
            keep[k] = 1;
        }
        int have = 0;
        double s = 0.0;
        for (int round = 0; round < 3; ++round) {
            double g[6];
            if (!fit_affine((const double (*)[4])nb, keep, nc, g)) break;
            for (size_t k = 0; k < nc; ++k) {
                const double rx = nb[k][2] - (g[0] + g[1] * nb[k][0] + g[2] * nb[k][1]);
                const double ry = nb[k][3] - (g[3] + g[4] * nb[k][0] + g[5] * nb[k][1]);
                res[k] = sqrt(rx * rx + ry * ry);
                sorted[k] = res[k];
            }
            qsort(sorted, nc, sizeof(double), cmp_f64);
            s = 1.4826 * sorted[nc / 2];
            memcpy(f, g, sizeof g);
            have = 1;
            if (round == 2) break;
            const double cut = fmax(3.0 * s, tol * 0.25);
            size_t kept = 0;
            for (size_t k = 0; k < nc; ++k) {
                keep[k] = res[k] <= cut;
                kept += keep[k];
            }
            if (kept < 6) break;
        }
        if (!have) continue;
        const double r = sqrt(f[0] * f[0] + f[3] * f[3]);
        double lim = 4.0 * s;
        const double lo_l = tol * floor_, hi_l = cap_ * tol;
        if (lim < lo_l) lim = lo_l;
        if (lim > hi_l) lim = hi_l;
        out[i] = r > lim;
        if (offs) {
            offs[2 * i] = r;
            offs[2 * i + 1] = lim;
        }
    }
done:
    free(px);
    free(py);
    free(dx);
    free(dy);
    free(tmp);
    free(idx);
    free(cand);
}

void orc_displacement_outliers(const uint64_t* m, size_t n, const double* kpa, const double* kpb,
                               double floor_, double cap_, uint8_t* out) {
    displacement_impl(m, n, kpa, kpb, floor_, cap_, out, NULL);
}

void orc_displacement_offsets(const uint64_t* m, size_t n, const double* kpa, const double* kpb,
                              double floor_, double cap_, uint8_t* out, double* offs) {
    for (size_t i = 0; i < 2 * n; ++i) offs[i] = NAN;
    displacement_impl(m, n, kpa, kpb, floor_, cap_, out, offs);
}

/* ----------------------------------------------------------- union-find */

/* UnionFind (tracks.rs:160-188): find with path compression, link keeps the
 * smaller root. */
static size_t uf_find(size_t* parent, size_t x) {
    size_t r = x;
    while (parent[r] != r) r = parent[r];
    while (parent[x] != r) {
        const size_t next = parent[x];
        parent[x] = r;
        x = next;
    }
    return r;
}
static size_t uf_link(size_t* parent, size_t ra, size_t rb) {
    const size_t lo = ra < rb ? ra : rb, hi = ra < rb ? rb : ra;
    parent[hi] = lo;
    return lo;
}

/* common_count / merge_sorted (tracks.rs:750-784) over sorted lists. */
static size_t common_count(const size_t* a, size_t na, const size_t* b, size_t nb) {
    size_t i = 0, j = 0, c = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) ++i;
        else if (a[i] > b[j]) ++j;
        else {
            ++c;
            ++i;
            ++j;
        }
    }
    return c;
}

/* The image list of each component root (sorted; empty = a lone node). */
typedef struct { size_t* v; size_t n; } ilist;

/* try_join (tracks.rs:692-724). */
static int try_join(size_t* parent, ilist* images, const size_t* node_image, size_t u, size_t v) {
    const size_t ru = uf_find(parent, u), rv = uf_find(parent, v);
    if (ru == rv) return 1;
    size_t su = node_image[ru], sv = node_image[rv];
    const size_t* iu = images[ru].n ? images[ru].v : &su;
    const size_t niu = images[ru].n ? images[ru].n : 1;
    const size_t* iv = images[rv].n ? images[rv].v : &sv;
    const size_t niv = images[rv].n ? images[rv].n : 1;
    if (common_count(iu, niu, iv, niv) > 0) return 0;
    size_t* mrg = malloc((niu + niv) * sizeof(size_t));
    size_t i = 0, j = 0, k = 0;
    while (i < niu && j < niv) mrg[k++] = iu[i] <= iv[j] ? iu[i++] : iv[j++];
    while (i < niu) mrg[k++] = iu[i++];
    while (j < niv) mrg[k++] = iv[j++];
    free(images[ru].v);
    free(images[rv].v);
    images[ru].v = images[rv].v = NULL;
    images[ru].n = images[rv].n = 0;
    const size_t r = uf_link(parent, ru, rv);
    images[r].v = mrg;
    images[r].n = k;
    return 1;
}

typedef struct { size_t a, b; } edge_t;
static int cmp_edge(const void* x, const void* y) {
    const edge_t* p = x;
    const edge_t* q = y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    return p->b < q->b ? -1 : (p->b > q->b);
}
typedef struct { size_t sup, u, v; } oedge_t;
static int cmp_oedge(const void* x, const void* y) { /* support descending, then (u, v) */
    const oedge_t* p = x;
    const oedge_t* q = y;
    if (p->sup != q->sup) return p->sup > q->sup ? -1 : 1;
    if (p->u != q->u) return p->u < q->u ? -1 : 1;
    return p->v < q->v ? -1 : (p->v > q->v);
}
typedef struct { size_t a, b, count, sup; } cpair_t;
static int cmp_cpair_strong(const void* x, const void* y) { /* (count, sup) desc, (a, b) asc */
    const cpair_t* p = x;
    const cpair_t* q = y;
    if (p->count != q->count) return p->count > q->count ? -1 : 1;
    if (p->sup != q->sup) return p->sup > q->sup ? -1 : 1;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    return p->b < q->b ? -1 : (p->b > q->b);
}
static int cmp_cpair_count(const void* x, const void* y) { /* count desc, (a, b) asc */
    const cpair_t* p = x;
    const cpair_t* q = y;
    if (p->count != q->count) return p->count > q->count ? -1 : 1;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    return p->b < q->b ? -1 : (p->b > q->b);
}
static int cmp_size(const void* x, const void* y) {
    const size_t a = *(const size_t*)x, b = *(const size_t*)y;
    return a < b ? -1 : (a > b);
}

/* component_pairs (tracks.rs:727-747): (smaller root, larger root, edges, 0),
 * ascending. Returns the count; out holds up to nz entries. */
static size_t component_pairs(const oedge_t* zero, size_t nz, size_t* parent, cpair_t* out,
                              edge_t* tmp) {
    size_t ng = 0;
    for (size_t k = 0; k < nz; ++k) {
        const size_t ru = uf_find(parent, zero[k].u), rv = uf_find(parent, zero[k].v);
        if (ru != rv) {
            tmp[ng].a = ru < rv ? ru : rv;
            tmp[ng].b = ru < rv ? rv : ru;
            ++ng;
        }
    }
    qsort(tmp, ng, sizeof(edge_t), cmp_edge);
    size_t nc = 0;
    for (size_t k = 0; k < ng; ++k) {
        if (nc && out[nc - 1].a == tmp[k].a && out[nc - 1].b == tmp[k].b) out[nc - 1].count++;
        else out[nc++] = (cpair_t){tmp[k].a, tmp[k].b, 1, 0};
    }
    return nc;
}

/* ------------------------------------------------------- select_tracks */

typedef struct {
    size_t len, span, image, feature;
    size_t t;
} tkey_t;
static int cmp_tkey_cap(const void* x, const void* y) { /* len desc, span desc, first asc */
    const tkey_t* p = x;
    const tkey_t* q = y;
    if (p->len != q->len) return p->len > q->len ? -1 : 1;
    if (p->span != q->span) return p->span > q->span ? -1 : 1;
    if (p->image != q->image) return p->image < q->image ? -1 : 1;
    return p->feature < q->feature ? -1 : (p->feature > q->feature);
}
static int cmp_tkey_first(const void* x, const void* y) {
    const tkey_t* p = x;
    const tkey_t* q = y;
    if (p->image != q->image) return p->image < q->image ? -1 : 1;
    return p->feature < q->feature ? -1 : (p->feature > q->feature);
}

/* select_tracks (tracks.rs:786-799) over tracks given as CSR of (image,
 * feature): writes the selected tracks' indices in output order, returns
 * their count. Track first observations are distinct, so both sorts are
 * total orders. */
long orc_select_tracks(const uint64_t* track_off, const uint64_t* obs, size_t n_tracks,
                       size_t max_tracks, uint64_t* order_out) {
    tkey_t* k = malloc((n_tracks + 1) * sizeof(tkey_t));
    if (!k) return -1;
    for (size_t t = 0; t < n_tracks; ++t) {
        const uint64_t a = track_off[t], b = track_off[t + 1];
        k[t].len = b - a;
        k[t].span = b > a ? obs[2 * (b - 1)] - obs[2 * a] : 0;
        k[t].image = b > a ? obs[2 * a] : 0;
        k[t].feature = b > a ? obs[2 * a + 1] : 0;
        k[t].t = t;
    }
    size_t n = n_tracks;
    if (n > max_tracks) {
        qsort(k, n, sizeof(tkey_t), cmp_tkey_cap);
        n = max_tracks;
    }
    qsort(k, n, sizeof(tkey_t), cmp_tkey_first);
    for (size_t t = 0; t < n; ++t) order_out[t] = k[t].t;
    free(k);
    return (long)n;
}

/* ------------------------------------------------------- build_tracks */

typedef struct { size_t root, node; } member_t;
static int cmp_member(const void* x, const void* y) {
    const member_t* p = x;
    const member_t* q = y;
    if (p->root != q->root) return p->root < q->root ? -1 : 1;
    return p->node < q->node ? -1 : (p->node > q->node);
}

/* build_tracks (tracks.rs:190-508). */
long orc_build_tracks(size_t n_pairs, const uint64_t* image_a, const uint64_t* image_b,
                      const uint64_t* m_off, const uint64_t* m_ab, size_t n_images,
                      const uint64_t* kp_off, const double* kp, const orc_track_cfg* cfg,
                      uint64_t* track_off, size_t track_cap, uint64_t* obs, size_t obs_cap,
                      orc_track_stats* st) {
    memset(st, 0, sizeof *st);
    const size_t n = (size_t)kp_off[n_images];
    size_t* node_image = malloc((n + 1) * sizeof(size_t));
    size_t max_m = 0;
    for (size_t p = 0; p < n_pairs; ++p)
        if (m_off[p + 1] - m_off[p] > max_m) max_m = m_off[p + 1] - m_off[p];
    size_t edge_cap = (size_t)m_off[n_pairs] + 1;
    edge_t* edges = malloc(edge_cap * sizeof(edge_t));
    uint64_t* valid = malloc((2 * max_m + 2) * sizeof(uint64_t));
    uint8_t* outl = malloc(max_m + 1);
    size_t* parent = malloc((n + 1) * sizeof(size_t));
    long ret = -1;
    if (!node_image || !edges || !valid || !outl || !parent) goto fail0;
    for (size_t i = 0; i < n_images; ++i)
        for (uint64_t x = kp_off[i]; x < kp_off[i + 1]; ++x) node_image[x] = i;
    /* tracks.rs:207-247: valid matches per pair, canonical direction, sorted,
     * deduplicated, displacement filter (Split), then the (u < v) edges */
    size_t ne = 0;
    for (size_t p = 0; p < n_pairs; ++p) {
        size_t nv = 0;
        const uint64_t ia = image_a[p], ib = image_b[p];
        for (uint64_t k = m_off[p]; k < m_off[p + 1]; ++k) {
            const uint64_t fa = m_ab[2 * k], fb = m_ab[2 * k + 1];
            const int ok = ia != ib && ia < n_images && ib < n_images &&
                           fa < kp_off[ia + 1] - kp_off[ia] && fb < kp_off[ib + 1] - kp_off[ib];
            if (!ok) {
                st->invalid_matches++;
                continue;
            }
            valid[2 * nv] = fa;
            valid[2 * nv + 1] = fb;
            ++nv;
        }
        if (ia > ib)
            for (size_t k = 0; k < nv; ++k) {
                const uint64_t t = valid[2 * k];
                valid[2 * k] = valid[2 * k + 1];
                valid[2 * k + 1] = t;
            }
        qsort(valid, nv, 2 * sizeof(uint64_t), cmp_edge);
        size_t nd = 0;
        for (size_t k = 0; k < nv; ++k)
            if (!nd || valid[2 * (nd - 1)] != valid[2 * k] || valid[2 * (nd - 1) + 1] != valid[2 * k + 1]) {
                valid[2 * nd] = valid[2 * k];
                valid[2 * nd + 1] = valid[2 * k + 1];
                ++nd;
            }
        const uint64_t lo = ia < ib ? ia : ib, hi = ia < ib ? ib : ia;
        if (cfg->policy == 1 && nd > 0)
            orc_displacement_outliers(valid, nd, kp + 2 * kp_off[lo], kp + 2 * kp_off[hi], DISP_FLOOR,
                                      DISP_CAP, outl);
        else
            memset(outl, 0, nd);
        for (size_t k = 0; k < nd; ++k) {
            if (outl[k]) {
                st->inconsistent++;
                continue;
            }
            const size_t u = kp_off[lo] + valid[2 * k], v = kp_off[hi] + valid[2 * k + 1];
            edges[ne].a = u < v ? u : v;
            edges[ne].b = u < v ? v : u;
            ++ne;
        }
    }
    qsort(edges, ne, sizeof(edge_t), cmp_edge);
    {
        size_t nd = 0;
        for (size_t k = 0; k < ne; ++k)
            if (!nd || edges[nd - 1].a != edges[k].a || edges[nd - 1].b != edges[k].b) edges[nd++] = edges[k];
        ne = nd;
    }
    st->edges = ne;
    for (size_t x = 0; x < n; ++x) parent[x] = x;

    if (cfg->policy == 0) { /* Drop (tracks.rs:251-258) */
        for (size_t k = 0; k < ne; ++k) {
            const size_t ru = uf_find(parent, edges[k].a), rv = uf_find(parent, edges[k].b);
            if (ru != rv) uf_link(parent, ru, rv);
        }
    } else { /* Split (tracks.rs:259-465) */
        size_t* start = calloc(n + 1, sizeof(size_t));
        size_t* fill = malloc((n + 1) * sizeof(size_t));
        size_t* flat = malloc((2 * ne + 1) * sizeof(size_t));
        size_t* fe = malloc((2 * ne + 1) * sizeof(size_t));
        size_t* sup = malloc((ne + 1) * sizeof(size_t));
        uint8_t* cut = malloc(ne + 1);
        size_t* newid = malloc((ne + 1) * sizeof(size_t));
        oedge_t* order = malloc((ne + 1) * sizeof(oedge_t));
        ilist* images = calloc(n + 1, sizeof(ilist));
        cpair_t* counted = malloc((ne + 1) * sizeof(cpair_t));
        edge_t* tmp = malloc((ne + 1) * sizeof(edge_t));
        size_t* cstart = malloc((n + 1) * sizeof(size_t));
        size_t* cfill = malloc((n + 1) * sizeof(size_t));
        size_t* cflat = malloc((2 * ne + 1) * sizeof(size_t));
        int okm = start && fill && flat && fe && sup && cut && newid && order && images && counted &&
                  tmp && cstart && cfill && cflat;
        if (okm) {
            for (size_t k = 0; k < ne; ++k) {
                start[edges[k].a + 1]++;
                start[edges[k].b + 1]++;
            }
            for (size_t i = 0; i < n; ++i) start[i + 1] += start[i];
            memcpy(fill, start, (n + 1) * sizeof(size_t));
            for (size_t e = 0; e < ne; ++e) {
                const size_t u = edges[e].a, v = edges[e].b;
                flat[fill[u]] = v;
                fe[fill[u]] = e;
                fill[u]++;
                flat[fill[v]] = u;
                fe[fill[v]] = e;
                fill[v]++;
            }
            for (int round = 0; round < VOTE_ROUNDS; ++round) {
                for (size_t e = 0; e < ne; ++e) {
                    const size_t u = edges[e].a, v = edges[e].b;
                    sup[e] = common_count(flat + start[u], start[u + 1] - start[u], flat + start[v],
                                          start[v + 1] - start[v]);
                }
                memset(cut, 0, ne);
                size_t cuts = 0;
                for (size_t x = 0; x < n; ++x) {
                    const size_t lo = start[x], hi = start[x + 1];
                    size_t i = lo;
                    while (i < hi) {
                        const size_t img = node_image[flat[i]];
                        size_t j = i + 1;
                        while (j < hi && node_image[flat[j]] == img) ++j;
                        if (j - i > 1) {
                            size_t best = 0, top = 0;
                            for (size_t q = i; q < j; ++q)
                                if (sup[fe[q]] > best) best = sup[fe[q]];
                            for (size_t q = i; q < j; ++q) top += sup[fe[q]] == best;
                            for (size_t q = i; q < j; ++q) {
                                const size_t e = fe[q];
                                const int drop = sup[e] < best ? best > 0 : (top > 1 && best > 0);
                                if (drop && !cut[e]) {
                                    cut[e] = 1;
                                    ++cuts;
                                }
                            }
                        }
                        i = j;
                    }
                }
                if (cuts == 0) break;
                st->conflicts += cuts;
                st->vote_cuts += cuts;
                size_t kk = 0;
                for (size_t e = 0; e < ne; ++e) newid[e] = cut[e] ? SIZE_MAX : kk++;
                size_t w = 0;
                for (size_t e = 0; e < ne; ++e)
                    if (!cut[e]) edges[w++] = edges[e];
                size_t wf = 0, r = 0;
                for (size_t x = 0; x < n; ++x) {
                    const size_t end = start[x + 1];
                    start[x] = wf;
                    while (r < end) {
                        if (!cut[fe[r]]) {
                            flat[wf] = flat[r];
                            fe[wf] = newid[fe[r]];
                            ++wf;
                        }
                        ++r;
                    }
                }
                start[n] = wf;
                ne = w;
            }
            for (size_t e = 0; e < ne; ++e) {
                const size_t u = edges[e].a, v = edges[e].b;
                order[e].sup = common_count(flat + start[u], start[u + 1] - start[u], flat + start[v],
                                            start[v + 1] - start[v]);
                order[e].u = u;
                order[e].v = v;
            }
            qsort(order, ne, sizeof(oedge_t), cmp_oedge);
            size_t split = ne;
            for (size_t e = 0; e < ne; ++e)
                if (order[e].sup == 0) {
                    split = e;
                    break;
                }
            for (size_t e = 0; e < split; ++e)
                if (!try_join(parent, images, node_image, order[e].u, order[e].v)) st->conflicts++;
            const oedge_t* zero = order + split;
            const size_t nz = ne - split;
            for (int round = 0; round <= 8; ++round) {
                size_t nc = component_pairs(zero, nz, parent, counted, tmp);
                if (nc == 0) break;
                memset(cstart, 0, (n + 1) * sizeof(size_t));
                for (size_t c = 0; c < nc; ++c) {
                    cstart[counted[c].a + 1]++;
                    cstart[counted[c].b + 1]++;
                }
                for (size_t i = 0; i < n; ++i) cstart[i + 1] += cstart[i];
                memcpy(cfill, cstart, (n + 1) * sizeof(size_t));
                for (size_t c = 0; c < nc; ++c) {
                    cflat[cfill[counted[c].a]++] = counted[c].b;
                    cflat[cfill[counted[c].b]++] = counted[c].a;
                }
                for (size_t i = 0; i < n; ++i)
                    if (cstart[i + 1] - cstart[i] > 1)
                        qsort(cflat + cstart[i], cstart[i + 1] - cstart[i], sizeof(size_t), cmp_size);
                for (size_t c = 0; c < nc; ++c) {
                    const size_t a = counted[c].a, b = counted[c].b;
                    counted[c].sup = common_count(cflat + cstart[a], cstart[a + 1] - cstart[a],
                                                  cflat + cstart[b], cstart[b + 1] - cstart[b]);
                }
                qsort(counted, nc, sizeof(cpair_t), cmp_cpair_strong);
                const int last = round == 8;
                size_t merged = 0;
                for (size_t c = 0; c < nc; ++c) {
                    const int strong = counted[c].count >= 2 || counted[c].sup > 0;
                    if (!strong && !last) continue;
                    if (try_join(parent, images, node_image, counted[c].a, counted[c].b)) ++merged;
                    else st->conflicts += counted[c].count;
                }
                if (last) break;
                if (merged == 0) {
                    const size_t nr = component_pairs(zero, nz, parent, counted, tmp);
                    qsort(counted, nr, sizeof(cpair_t), cmp_cpair_count);
                    for (size_t c = 0; c < nr; ++c)
                        if (!try_join(parent, images, node_image, counted[c].a, counted[c].b))
                            st->conflicts += counted[c].count;
                    break;
                }
            }
            for (size_t i = 0; i < n; ++i) free(images[i].v);
        }
        free(start);
        free(fill);
        free(flat);
        free(fe);
        free(sup);
        free(cut);
        free(newid);
        free(order);
        free(images);
        free(counted);
        free(tmp);
        free(cstart);
        free(cfill);
        free(cflat);
        if (!okm) goto fail0;
    }

    /* tracks.rs:467-508: components of the remaining edges' end points */
    {
        member_t* mem = malloc((2 * ne + 1) * sizeof(member_t));
        tkey_t* keys = NULL;
        uint64_t* toff = NULL;
        uint64_t* tobs = NULL;
        if (!mem) goto fail0;
        size_t nm = 0;
        for (size_t e = 0; e < ne; ++e) {
            mem[nm++] = (member_t){uf_find(parent, edges[e].a), edges[e].a};
            mem[nm++] = (member_t){uf_find(parent, edges[e].b), edges[e].b};
        }
        qsort(mem, nm, sizeof(member_t), cmp_member);
        const size_t min_len = cfg->min_length > 2 ? (size_t)cfg->min_length : 2;
        toff = malloc((nm / 2 + 2) * sizeof(uint64_t));
        tobs = malloc((nm + 1) * 2 * sizeof(uint64_t));
        if (!toff || !tobs) {
            free(mem);
            free(toff);
            free(tobs);
            goto fail0;
        }
        size_t nt = 0, no = 0;
        toff[0] = 0;
        size_t g = 0;
        while (g < nm) {
            size_t h = g;
            while (h < nm && mem[h].root == mem[g].root) ++h;
            /* nodes of this component, sorted (they are) and deduplicated */
            size_t cnt = 0, first = no;
            int conflict = 0;
            for (size_t q = g; q < h; ++q) {
                if (q > g && mem[q].node == mem[q - 1].node) continue;
                const size_t x = mem[q].node;
                if (cnt && node_image[x] == tobs[2 * (no - 1)]) conflict = 1;
                tobs[2 * no] = node_image[x];
                tobs[2 * no + 1] = x - kp_off[node_image[x]];
                ++no;
                ++cnt;
            }
            if (conflict) {
                st->conflicts++;
                no = first;
            } else if (cnt < min_len) {
                st->too_short++;
                no = first;
            } else {
                toff[++nt] = no;
            }
            g = h;
        }
        free(mem);
        /* select_tracks */
        keys = malloc((nt + 1) * sizeof(tkey_t));
        uint64_t* sel = malloc((nt + 1) * sizeof(uint64_t));
        if (!keys || !sel) {
            free(keys);
            free(sel);
            free(toff);
            free(tobs);
            goto fail0;
        }
        const long ns = orc_select_tracks(toff, tobs, nt, (size_t)cfg->max_tracks, sel);
        st->truncated = nt - (size_t)ns;
        st->tracks = (size_t)ns;
        size_t total = 0;
        for (long t = 0; t < ns; ++t) total += toff[sel[t] + 1] - toff[sel[t]];
        if ((size_t)ns + 1 > track_cap || total > obs_cap) {
            ret = -1;
        } else {
            size_t w = 0;
            track_off[0] = 0;
            for (long t = 0; t < ns; ++t) {
                for (uint64_t q = toff[sel[t]]; q < toff[sel[t] + 1]; ++q) {
                    obs[2 * w] = tobs[2 * q];
                    obs[2 * w + 1] = tobs[2 * q + 1];
                    ++w;
                }
                track_off[t + 1] = w;
            }
            ret = ns;
        }
        free(keys);
        free(sel);
        free(toff);
        free(tobs);
    }
fail0:
    free(node_image);
    free(edges);
    free(valid);
    free(outl);
    free(parent);
    return ret;
}
