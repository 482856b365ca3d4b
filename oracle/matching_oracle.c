/*
 * CPU oracle, descriptor-matching half. TEST INFRASTRUCTURE ONLY (see
 * oracle.h): the checker for the HIP path, never the product.
 *
 * Restates crates/core/src/matching.rs:14-40 (candidate_pairs), :41-142
 * (PairSchedule, scheduled_pairs), :144-215 (sq_dist, ratio_match) and the
 * test helpers :1141-1154 (Lcg), :1580-1656 (feature, ratio_match_reference,
 * match_fixture), :1867-1878 (pairs_hash), :2345-2357 (sq_dist_spec); and
 * crates/core/src/tracks.rs:56-69 (verified_matches).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* matching.rs:1143-1149: LCG step, top 53 bits / 2^53. */
double orc_lcg_next(orc_lcg* g) {
    g->s = g->s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(g->s >> 11) / (double)(1ULL << 53);
}

/* matching.rs:1150-1153: Box-Muller, u1 clamped at 1e-300. */
double orc_lcg_gauss(orc_lcg* g) {
    double u1 = orc_lcg_next(g);
    if (!(u1 >= 1e-300)) u1 = 1e-300; /* f64::max(1e-300): NaN-free here */
    double u2 = orc_lcg_next(g);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* matching.rs:1580-1595: 128 uniforms cast to f32, divided by the f32 norm
 * (sequential f32 sum of squares, then sqrtf). */
void orc_feature(orc_lcg* g, float* desc) {
    for (int k = 0; k < ORC_DESC_LEN; ++k) desc[k] = (float)orc_lcg_next(g);
    float s = -0.0f;
    for (int k = 0; k < ORC_DESC_LEN; ++k) s = s + desc[k] * desc[k];
    float n = sqrtf(s);
    for (int k = 0; k < ORC_DESC_LEN; ++k) desc[k] = desc[k] / n;
}

size_t orc_match_fixture_nb(size_t na, size_t extra) {
    size_t nb = extra;
    for (size_t i = 0; i < na; ++i) {
        if (i % 3 == 2) continue;
        nb += 1 + (i % 7 == 0 ? 1 : 0);
    }
    return nb;
}

/* matching.rs:1638-1656: b = noisy copies (0.03 f32 * gauss) of a, twins for
 * i % 7 == 0, skipping i % 3 == 2, then `extra` unrelated features. */
long orc_match_fixture(uint64_t seed, size_t na, size_t extra, float* a, float* b, size_t b_cap) {
    size_t nb = orc_match_fixture_nb(na, extra);
    if (nb > b_cap) return -1;
    orc_lcg g = {seed};
    for (size_t i = 0; i < na; ++i) orc_feature(&g, a + i * ORC_DESC_LEN);
    size_t r = 0;
    for (size_t i = 0; i < na; ++i) {
        if (i % 3 == 2) continue;
        size_t copies = 1 + (i % 7 == 0 ? 1 : 0);
        for (size_t c = 0; c < copies; ++c) {
            float* h = b + r * ORC_DESC_LEN;
            memcpy(h, a + i * ORC_DESC_LEN, sizeof(float) * ORC_DESC_LEN);
            for (int k = 0; k < ORC_DESC_LEN; ++k) h[k] += 0.03f * (float)orc_lcg_gauss(&g);
            ++r;
        }
    }
    for (size_t e = 0; e < extra; ++e) orc_feature(&g, b + (r++) * ORC_DESC_LEN);
    return (long)nb;
}

/* matching.rs:144-155: lane k accumulates components 8i+k (t = p - r;
 * acc += t * t, separately rounded), lanes summed 0..7 left to right. */
float orc_sq_dist(const float* p, const float* r) {
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < ORC_DESC_LEN / 8; ++c) {
        for (int k = 0; k < 8; ++k) {
            float t = p[8 * c + k] - r[8 * c + k];
            acc[k] += t * t;
        }
    }
    /* Iterator::sum for f32 folds from -0.0; -0.0 + x == x for x >= +0. */
    float s = -0.0f;
    for (int k = 0; k < 8; ++k) s = s + acc[k];
    return s;
}

/* matching.rs:2345-2357: the written specification of sq_dist. */
float orc_sq_dist_spec(const float* p, const float* r) {
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < ORC_DESC_LEN; ++i) {
        float t = p[i] - r[i];
        acc[i % 8] += t * t;
    }
    float s = 0.0f;
    for (int k = 0; k < 8; ++k) s += acc[k];
    return s;
}

/* matching.rs:2372-2376: the sequential sum the boundary test contrasts with. */
float orc_sq_dist_seq(const float* p, const float* r) {
    float s = 0.0f;
    for (int i = 0; i < ORC_DESC_LEN; ++i) s = s + (p[i] - r[i]) * (p[i] - r[i]);
    return s;
}

/* matching.rs:167-203: per 64-row block of a, row (j1, d1, d2) over j
 * ascending and column (d, i) with strict '<'; blocks merged in order. */
void orc_ratio_match_state(const float* a, size_t na, const float* b, size_t nb,
                           uint64_t* row_j1, float* row_d1, float* row_d2, float* col_d,
                           uint64_t* col_i) {
    for (size_t j = 0; j < nb; ++j) { col_d[j] = INFINITY; col_i[j] = UINT64_MAX; }
    float* blk_d = (float*)malloc(sizeof(float) * (nb ? nb : 1));
    uint64_t* blk_i = (uint64_t*)malloc(sizeof(uint64_t) * (nb ? nb : 1));
    for (size_t b0 = 0; b0 < na; b0 += ORC_MATCH_BLOCK) {
        size_t len = na - b0 < ORC_MATCH_BLOCK ? na - b0 : ORC_MATCH_BLOCK;
        for (size_t r = 0; r < len; ++r) {
            row_j1[b0 + r] = UINT64_MAX; row_d1[b0 + r] = INFINITY; row_d2[b0 + r] = INFINITY;
        }
        for (size_t j = 0; j < nb; ++j) { blk_d[j] = INFINITY; blk_i[j] = UINT64_MAX; }
        for (size_t j = 0; j < nb; ++j) {
            for (size_t r = 0; r < len; ++r) {
                size_t i = b0 + r;
                float d = orc_sq_dist(a + i * ORC_DESC_LEN, b + j * ORC_DESC_LEN);
                if (d < row_d1[i]) {
                    row_d2[i] = row_d1[i]; row_d1[i] = d; row_j1[i] = j;
                } else if (d < row_d2[i]) {
                    row_d2[i] = d;
                }
                if (d < blk_d[j]) { blk_d[j] = d; blk_i[j] = i; }
            }
        }
        /* matching.rs:196-203: merge in block order, strict '<'. */
        for (size_t j = 0; j < nb; ++j)
            if (blk_d[j] < col_d[j]) { col_d[j] = blk_d[j]; col_i[j] = blk_i[j]; }
    }
    free(blk_d);
    free(blk_i);
}

/* matching.rs:204-214: accept iff j1 exists, d2 finite, d1 < ratio*ratio*d2
 * (f32), and (mutual) col_best[j1].i == i. */
size_t orc_ratio_match(const float* a, size_t na, const float* b, size_t nb, float ratio,
                       int mutual, uint64_t* out_ij) {
    if (na == 0) return 0;
    uint64_t* j1 = (uint64_t*)malloc(sizeof(uint64_t) * na);
    float* d1 = (float*)malloc(sizeof(float) * na);
    float* d2 = (float*)malloc(sizeof(float) * na);
    float* cd = (float*)malloc(sizeof(float) * (nb ? nb : 1));
    uint64_t* ci = (uint64_t*)malloc(sizeof(uint64_t) * (nb ? nb : 1));
    orc_ratio_match_state(a, na, b, nb, j1, d1, d2, cd, ci);
    size_t n = 0;
    float r2 = ratio * ratio;
    for (size_t i = 0; i < na; ++i) {
        if (j1[i] == UINT64_MAX || !isfinite(d2[i]) || d1[i] >= r2 * d2[i]) continue;
        if (mutual && ci[j1[i]] != i) continue;
        out_ij[2 * n] = i;
        out_ij[2 * n + 1] = j1[i];
        ++n;
    }
    free(j1); free(d1); free(d2); free(cd); free(ci);
    return n;
}

/* matching.rs:1604-1620: nearest/second over `set`; sum of powi(2) terms
 * in sequential order. */
static int brute_nearest(const float* q, const float* set, size_t ns, uint64_t* j, float* d1,
                         float* d2) {
    uint64_t bj = UINT64_MAX;
    float b1 = INFINITY, b2 = INFINITY;
    for (size_t k = 0; k < ns; ++k) {
        const float* f = set + k * ORC_DESC_LEN;
        float d = -0.0f;
        for (int c = 0; c < ORC_DESC_LEN; ++c) {
            float t = q[c] - f[c];
            d = d + t * t;
        }
        if (d < b1) { b2 = b1; b1 = d; bj = k; }
        else if (d < b2) b2 = d;
    }
    *j = bj; *d1 = b1; *d2 = b2;
    return bj != UINT64_MAX;
}

/* matching.rs:1598-1635. */
size_t orc_ratio_match_brute(const float* a, size_t na, const float* b, size_t nb, float ratio,
                             int mutual, uint64_t* out_ij) {
    size_t n = 0;
    for (size_t i = 0; i < na; ++i) {
        uint64_t j; float d1, d2;
        if (!brute_nearest(a + i * ORC_DESC_LEN, b, nb, &j, &d1, &d2)) continue;
        if (!isfinite(d2) || d1 >= ratio * ratio * d2) continue;
        if (mutual) {
            uint64_t ii; float e1, e2;
            if (!brute_nearest(b + j * ORC_DESC_LEN, a, na, &ii, &e1, &e2) || ii != i) continue;
        }
        out_ij[2 * n] = i;
        out_ij[2 * n + 1] = j;
        ++n;
    }
    return n;
}

/* matching.rs:1867-1878. */
uint64_t orc_pairs_hash(const uint64_t* ij, size_t n) {
    uint64_t h = 0xcbf29ce484222325ULL;
    for (size_t k = 0; k < 2 * n; ++k) {
        uint64_t v = ij[k];
        for (int byte = 0; byte < 8; ++byte) {
            h ^= (v >> (8 * byte)) & 0xff;
            h *= 0x100000001b3ULL;
        }
    }
    return h;
}

/* matching.rs:18-40: same camera -> 1 <= d <= temporal or (d a power of two
 * and d <= pow2_max); other camera -> d <= cross. i ascending, j > i. */
size_t orc_candidate_pairs(const uint64_t* views, size_t n, uint64_t temporal, uint64_t cross,
                           uint64_t pow2_max, uint64_t* out, size_t cap) {
    size_t cnt = 0;
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = i + 1; j < n; ++j) {
            uint64_t ca = views[2 * i], pa = views[2 * i + 1];
            uint64_t cb = views[2 * j], pb = views[2 * j + 1];
            uint64_t d = pa > pb ? pa - pb : pb - pa;
            int ok;
            if (ca == cb) {
                int pow2 = d != 0 && (d & (d - 1)) == 0;
                ok = d >= 1 && (d <= temporal || (pow2 && d <= pow2_max));
            } else {
                ok = d <= cross;
            }
            if (ok) {
                if (out && cnt < cap) { out[2 * cnt] = i; out[2 * cnt + 1] = j; }
                ++cnt;
            }
        }
    }
    return cnt;
}

/* matching.rs:41-44: camera numbers (synth::CamId::ALL order). */
#define ORC_CAM_FRONT 0
#define ORC_CAM_RIGHT 1
#define ORC_CAM_LEFT 2

static int cmp_pair(const void* x, const void* y) {
    const uint64_t* a = (const uint64_t*)x;
    const uint64_t* b = (const uint64_t*)y;
    if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
    if (a[1] != b[1]) return a[1] < b[1] ? -1 : 1;
    return 0;
}

/* matching.rs:103-142: every ordered (i, j), i != j, tested by the schedule;
 * accepted ones pushed as (min, max); then sort_unstable and dedup. */
size_t orc_scheduled_pairs(const uint64_t* views, size_t n, const orc_pair_schedule* sch,
                           uint64_t* out, size_t cap) {
    size_t m = 0, mcap = 64;
    uint64_t* v = (uint64_t*)malloc(sizeof(uint64_t) * 2 * mcap);
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < n; ++j) {
            if (i == j) continue;
            uint64_t ca = views[2 * i], pa = views[2 * i + 1];
            uint64_t cb = views[2 * j], pb = views[2 * j + 1];
            uint64_t d = pa > pb ? pa - pb : pb - pa;
            int ok;
            if (ca == cb) {
                int pow2 = d != 0 && (d & (d - 1)) == 0; /* usize::is_power_of_two */
                ok = j > i && d >= 1 && (d <= sch->temporal || (pow2 && d <= sch->pow2_max));
            } else if (sch->cross_kind == 0) {
                ok = j > i && d <= sch->range;
            } else {
                int has_lo = 0;
                uint64_t lo = 0;
                if (ca == ORC_CAM_FRONT && cb == ORC_CAM_RIGHT) { has_lo = 1; lo = sch->right_min; }
                if (ca == ORC_CAM_FRONT && cb == ORC_CAM_LEFT) { has_lo = 1; lo = sch->left_min; }
                uint64_t step = sch->step > 1 ? sch->step : 1; /* step.max(1) */
                ok = has_lo && pb >= pa + lo && pb <= pa + sch->max && (pb - pa - lo) % step == 0;
            }
            if (ok) {
                if (m == mcap) {
                    mcap *= 2;
                    v = (uint64_t*)realloc(v, sizeof(uint64_t) * 2 * mcap);
                }
                v[2 * m] = i < j ? i : j;
                v[2 * m + 1] = i < j ? j : i;
                m++;
            }
        }
    }
    qsort(v, m, sizeof(uint64_t) * 2, cmp_pair);
    size_t k = 0;
    for (size_t e = 0; e < m; ++e) {
        if (k > 0 && v[2 * (k - 1)] == v[2 * e] && v[2 * (k - 1) + 1] == v[2 * e + 1]) continue;
        v[2 * k] = v[2 * e];
        v[2 * k + 1] = v[2 * e + 1];
        k++;
    }
    if (out)
        for (size_t e = 0; e < k && e < cap; ++e) {
            out[2 * e] = v[2 * e];
            out[2 * e + 1] = v[2 * e + 1];
        }
    free(v);
    return k;
}

/* tracks.rs:56-69. */
long orc_verified_matches(const uint64_t* ij, size_t n_matches, const uint8_t* inliers,
                          size_t n_inliers, uint64_t* out) {
    if (n_matches != n_inliers) return -1;
    long k = 0;
    for (size_t e = 0; e < n_matches; ++e) {
        if (!inliers[e]) continue;
        out[2 * k] = ij[2 * e];
        out[2 * k + 1] = ij[2 * e + 1];
        k++;
    }
    return k;
}
