/*
 * CPU oracle for the pairwise matching path of skylens-stream
 * (crates/core/src/matching.rs). TEST INFRASTRUCTURE ONLY: this library is
 * the checker that the HIP path is compared against. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path (skylens-stream-rs_amd/) never links or calls it.
 *
 * Every function restates the reference statement by statement and cites
 * the file:line it follows. Build flags: -O2 -ffp-contract=off, no fast-math
 * (Rust never contracts a*b+c into an FMA).
 */
#ifndef SKYLENS_ORACLE_H
#define SKYLENS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_DESC_LEN 128
#define ORC_MATCH_BLOCK 64

/* ---- deterministic test generators (matching.rs:1141-1154) ---- */
typedef struct { uint64_t s; } orc_lcg;
double orc_lcg_next(orc_lcg* g);
double orc_lcg_gauss(orc_lcg* g);

/* feature(): 128 uniforms, L2-normalised in f32 (matching.rs:1580-1595). */
void orc_feature(orc_lcg* g, float* desc);
/* match_fixture(seed, na, extra) (matching.rs:1638-1656). b_cap rows of
 * storage must be provided; returns the row count of b (-1 if b_cap short). */
long orc_match_fixture(uint64_t seed, size_t na, size_t extra, float* a, float* b, size_t b_cap);
size_t orc_match_fixture_nb(size_t na, size_t extra);

/* ---- descriptor matching ---- */
float orc_sq_dist(const float* p, const float* r);            /* matching.rs:144-155 */
float orc_sq_dist_spec(const float* p, const float* r);       /* matching.rs:2345-2357 */
float orc_sq_dist_seq(const float* p, const float* r);        /* matching.rs:2372-2376 */
/* ratio_match (matching.rs:157-215). out_ij holds 2*cap uint64 (i,j) pairs,
 * returns the match count (always <= na). */
size_t orc_ratio_match(const float* a, size_t na, const float* b, size_t nb, float ratio,
                       int mutual, uint64_t* out_ij);
/* ratio_match_reference: the brute-force form (matching.rs:1598-1635). */
size_t orc_ratio_match_brute(const float* a, size_t na, const float* b, size_t nb, float ratio,
                             int mutual, uint64_t* out_ij);
/* Exact row state and column winner of ratio_match, exposed so tests can
 * compare every intermediate: row_j1/row_d1/row_d2 (na each), col_d/col_i (nb). */
void orc_ratio_match_state(const float* a, size_t na, const float* b, size_t nb,
                           uint64_t* row_j1, float* row_d1, float* row_d2, float* col_d,
                           uint64_t* col_i);
/* FNV-1a 64 over (i, j) as u64 LE (matching.rs:1867-1878). */
uint64_t orc_pairs_hash(const uint64_t* ij, size_t n);

/* candidate_pairs (matching.rs:14-40). views: n (cam, pos) pairs as uint64.
 * Returns the pair count; writes pairs if out != NULL (cap pairs). */
size_t orc_candidate_pairs(const uint64_t* views, size_t n, uint64_t temporal, uint64_t cross,
                           uint64_t pow2_max, uint64_t* out, size_t cap);

/* PairSchedule / CrossSchedule (matching.rs:47-99). cross_kind 0 = Spec
 * { range }, 1 = Formation { right_min, left_min, max, step }. */
typedef struct {
    uint64_t temporal, pow2_max;
    int32_t cross_kind, pad;
    uint64_t range;
    uint64_t right_min, left_min, max, step;
} orc_pair_schedule;
/* scheduled_pairs (matching.rs:103-142): returns the pair count, writes up to
 * cap pairs (2 uint64 each) when out != NULL. */
size_t orc_scheduled_pairs(const uint64_t* views, size_t n, const orc_pair_schedule* sch,
                           uint64_t* out, size_t cap);
/* verified_matches (tracks.rs:56-69): 0 and the kept (i, j) in order, or -1
 * on a length mismatch. */
long orc_verified_matches(const uint64_t* ij, size_t n_matches, const uint8_t* inliers,
                          size_t n_inliers, uint64_t* out);

/* ---- fundamental-matrix RANSAC (matching.rs:217-1104) ---- */
typedef struct {
    double threshold_px;
    double min_inlier_ratio;
    uint64_t max_iters;
    double confidence;
    uint64_t seed;
    uint64_t min_inliers;
    double max_chance_prob;
} orc_ransac_cfg;

void orc_ransac_cfg_default(orc_ransac_cfg* c);  /* matching.rs:781-793 */

typedef struct {
    int has_selection;
    int model;          /* 0 = Fundamental, 1 = Homography */
    double gric_f, gric_h;
    uint64_t parallax, parallax_needed;
    double sigma_px;
} orc_model_sel;

uint64_t orc_adaptive_iterations(double w, int32_t sample, double confidence, uint64_t max_iters);
double orc_binomial_tail(uint64_t n, uint64_t k, double p);
double orc_chance_inlier_probability(uint64_t n, uint64_t cnt, double p);
double orc_epipolar_band_probability(const double* x2, size_t n, double th_px);
uint64_t orc_parallax_needed(uint64_t n);
/* F is written row-major (F[3*r+c]). x1, x2: n x 2 doubles. Return 1 = Some. */
int orc_fundamental_8pt(const double* x1, const double* x2, size_t n, double* F);
double orc_sampson_error(const double* F, const double* p, const double* q);
void orc_sampson_residual_jacobian(const double* g, double s1, double s2, const double* a,
                                   const double* b, double* r, double* j9);
void orc_normalizer(const double* p, size_t n, double* T);
int orc_refine_sampson(const double* F, const double* x1, const double* x2, size_t n, size_t iters,
                       double* out, uint64_t* taken, int* stop);
uint64_t orc_homography_support(const double* x1, const double* x2, size_t n, double th_px);
/* homography_dlt (matching.rs:251-266) of n correspondences, and the sampled
 * hypotheses' closed-form four-point solution; H row-major, unit Frobenius
 * norm. Return 0 for no homography. */
int orc_homography_dlt(const double* x1, const double* x2, size_t n, double* H);
int orc_homography_4pt(const double* x1, const double* x2, double* H);
int orc_select_two_view_model(const double* x1, const double* x2, size_t n, const double* F,
                              double sigma_px, orc_model_sel* out);
/* Returns 1 = Some(FundamentalFit). F row-major, inliers n bytes. */
int orc_ransac_fundamental(const double* x1, const double* x2, size_t n, const orc_ransac_cfg* cfg,
                           double* F, uint8_t* inliers, orc_model_sel* sel);
/* Diagnostics of the last orc_ransac_fundamental call on this thread:
 * hypotheses drawn, LO runs, LM calls. */
/* LM iterations and cost evaluations of the last orc_ransac_fundamental call. */
void orc_lm_work(uint64_t* iters, uint64_t* evals);
void orc_ransac_last_stats(uint64_t* iters, uint64_t* lo_runs, uint64_t* lm_calls,
                           uint64_t* lm_converged, uint64_t* lm_stalled_at_start);

/* Trace of the hypothesis loop (matching.rs:967-1036), for the tests that
 * pin the loop's semantics against independent restatements. One record per
 * iteration: the sample, whether fundamental_8pt gave Some, the minimal
 * sample's F and inlier count, the best count before the iteration, whether
 * local optimisation ran and its six stages, the LO result, the iteration
 * limit after the iteration, and the pool (count and origin iteration of each
 * entry) after it. lo_masks (cap x n bytes) receives each LO result's mask. */
typedef struct {
    uint64_t it;
    uint64_t idx[8];
    int32_t ok, lo;
    uint64_t raw_cnt, lo_cnt, best_before, iters_after;
    double F[9], F_lo[9];
    int32_t npool, pad;
    uint64_t pool_cnt[5], pool_origin[5];
    uint64_t stage_ns[6];
    int32_t stage_ok[6];
    uint64_t stage_cnt[6];
    double stage_F[6][9];
} orc_trace_rec;
/* Starts recording into recs (at most cap records) for the next
 * orc_ransac_fundamental call on this thread; n = its correspondence count. */
void orc_ransac_trace_begin(orc_trace_rec* recs, size_t cap, uint8_t* lo_masks, size_t n);
/* Stops recording; returns the number of records written. */
size_t orc_ransac_trace_end(void);

/* Eigen-solve of the symmetric 9x9 A used by the 8-point and DLT solvers
 * (stands in for nalgebra SymmetricEigen, matching.rs:539-541,
 * two_view.rs:638-639): the nine eigenvalues and a unit eigenvector of the
 * first smallest one. */
void orc_eigen_min9(const double* A, double* evals, double* evec);
/* The 8-point solver's eigen step (matching.rs:539-546): 1 if the normal
 * matrix is degenerate (second smallest eigenvalue at most 1e-10 of the
 * largest, Gershgorin-bounded), else 0 and a unit eigenvector of the smallest
 * eigenvalue by inverse iteration. */
int orc_eigen_min_inv9(const double* A, double* evec);

/* ---- test-geometry generators (matching.rs:1165-1220, camera.rs:41-52,123-159) ---- */
/* two_cameras(): K from_hfov(960,540,70deg), poses as in matching.rs:1165-1178.
 * cam: 9 (R row-major) + 3 (t) + 4 (fx, fy, cx, cy) + 2 (w, h) = 18 doubles. */
void orc_two_cameras(double* c1, double* c2);
int orc_project(const double* cam, const double* X, double* px);
void orc_fundamental_from_cameras(const double* c1, const double* c2, double* F);
/* correspondences(n, sigma, out, seed) (matching.rs:1182-1220). */
void orc_correspondences(size_t n, double sigma, double out, uint64_t seed, double* x1, double* x2,
                         uint8_t* truth);
/* planar_correspondences (matching.rs:1825-1845). */
void orc_planar_correspondences(size_t n, double sigma, uint64_t seed, double* x1, double* x2);
/* ground_with_buildings (matching.rs:1925-1952). */
void orc_ground_with_buildings(size_t n, double bld, double sigma, uint64_t seed, double* x1,
                               double* x2);
/* noisy_lines (matching.rs:1792-1806): 60 points per image. */
void orc_noisy_lines(double sigma, uint64_t seed, double* l1, double* l2);
/* line_plus_general (matching.rs:2201-2230): x1/x2 hold 60+k, h1/h2 hold 100. */
void orc_line_plus_general(size_t k, double sigma, uint64_t seed, double* x1, double* x2,
                           double* h1, double* h2);

/* ---- feature detector (features.rs:55-781, features_oracle.c) ---- */
/* DetectorConfig (features.rs:259-289); usize fields as uint64. */
typedef struct {
    uint64_t octaves, scales;
    float sigma0, contrast, edge_ratio;
    uint32_t upscale_fill;
    uint64_t max_features, upscale_below;
} orc_detector_cfg;
/* Keypoint (features.rs:295-304). */
typedef struct {
    float x, y, sigma, response, angle;
} orc_keypoint;
void orc_detector_cfg_default(orc_detector_cfg* c);
/* detect_and_describe (features.rs:594-648): writes min(total, cap)
 * features, returns the total (or -1 on allocation failure). */
long orc_detect_and_describe(const float* img, size_t w, size_t h, const orc_detector_cfg* cfg,
                             orc_keypoint* kps, float* desc, size_t cap);
/* gaussian_blur (features.rs:214-256); taps of it (returns r, -1 if cap short). */
int orc_gaussian_blur(const float* img, size_t w, size_t h, float sigma, float* out);
long orc_gaussian_taps(float sigma, float* k, size_t cap);
void orc_upsample2(const float* img, size_t w, size_t h, float* out);   /* features.rs:55-81 */
void orc_downsample(const float* img, size_t w, size_t h, float* out);  /* :200-210 */
/* dominant_orientations (features.rs:314-323), describe (:509-519). */
size_t orc_dominant_orientations(const float* img, size_t w, size_t h, float x, float y,
                                 float sigma, float* out);
void orc_describe(const float* img, size_t w, size_t h, float x, float y, float sigma,
                  float angle, float* desc);
/* f32 transcendentals evaluated in double (see features_oracle.c) and their
 * double kernels. */
float orc_expf(float x);
float orc_atan2f(float y, float x);
void orc_sincosf(float a, float* s, float* c);
float orc_powf(float b, float y);
double orc_xexp(double x);
double orc_xlog(double x);
double orc_xatan2(double y, double x);
void orc_xsincos(double a, double* s, double* c);

/* ---- multi-view tracks (tracks.rs:127-800) ---- */
/* TrackConfig (tracks.rs:109-125): policy 0 = Drop, 1 = Split. */
typedef struct {
    int32_t policy, pad;
    uint64_t min_length, max_tracks;
} orc_track_cfg;
/* TrackStats (tracks.rs:145-155). */
typedef struct {
    uint64_t edges, invalid_matches, conflicts, vote_cuts, too_short, inconsistent, truncated,
        tracks;
} orc_track_stats;
/* build_tracks (tracks.rs:190-508). Pair p: images image_a[p], image_b[p],
 * matches (fa, fb) = m_ab[2 k], m_ab[2 k + 1] for k in [m_off[p], m_off[p+1]).
 * Image i's keypoints: kp[2 x], kp[2 x + 1] for x in [kp_off[i], kp_off[i+1]).
 * Tracks out as CSR: track t's observations (image, feature) = obs[2 q],
 * obs[2 q + 1] for q in [track_off[t], track_off[t+1]). Returns the track
 * count, or -1 if track_cap (entries of track_off) or obs_cap (observations)
 * is short or an allocation fails. */
long orc_build_tracks(size_t n_pairs, const uint64_t* image_a, const uint64_t* image_b,
                      const uint64_t* m_off, const uint64_t* m_ab, size_t n_images,
                      const uint64_t* kp_off, const double* kp, const orc_track_cfg* cfg,
                      uint64_t* track_off, size_t track_cap, uint64_t* obs, size_t obs_cap,
                      orc_track_stats* st);
/* displacement_outliers_with_floor (tracks.rs:519-670) of n deduplicated
 * (fa, fb) of one pair: out[k] = 1 for a flagged match. */
void orc_displacement_outliers(const uint64_t* m, size_t n, const double* kpa, const double* kpb,
                               double floor_, double cap_, uint8_t* out);
/* Tests only: the flags and each match's fitted offset r and threshold (NAN where no fit ran). */
void orc_displacement_offsets(const uint64_t* m, size_t n, const double* kpa, const double* kpb,
                              double floor_, double cap_, uint8_t* out, double* offs);
/* Test knob: the order of the displacement filter's 24 kept candidates (0-3). */
void orc_set_displacement_order(int mode);
/* select_tracks (tracks.rs:786-799): the selected tracks' indices in output
 * order; returns their count (-1 on allocation failure). */
long orc_select_tracks(const uint64_t* track_off, const uint64_t* obs, size_t n_tracks,
                       size_t max_tracks, uint64_t* order_out);

#ifdef __cplusplus
}
#endif
#endif
