/*
 * skylens_hip.h — C-ABI of the MI355X (gfx950) pairwise-matching hot path of
 * skylens-stream: candidate_pairs -> ratio_match -> ransac_fundamental.
 *
 * Every entry point is plain C: pointers, sizes, status codes. No C++ types,
 * no exceptions cross this boundary. A Rust `skylens-hip` crate, a C++
 * caller or ctypes binds these symbols directly.
 *
 * Status codes: SKH_OK (0) on success; negative on error. Degenerate inputs
 * are NOT errors: they produce the reference's empty list / None
 * (matching.rs:946-949, :1680-1683, :1747-1773).
 *
 * Reference interfaces replaced (crates/core/src/matching.rs):
 *   candidate_pairs   :18-40   -> skh_candidate_pairs
 *   ratio_match       :167-215 -> skh_ratio_match / skh_ratio_match_dev
 *   ransac_fundamental:941-1104 -> skh_ransac_fundamental
 *   per-pair driver   benches/pipeline.rs:219-305 -> skh_plan_create / skh_plan_run /
 *                                                      skh_plan_results
 *   detect_and_describe  features.rs:594-648 -> skh_detect_and_describe
 *   build_tracks      tracks.rs:190-508 -> skh_build_tracks
 */
#ifndef SKYLENS_HIP_H
#define SKYLENS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SKH_OK 0
#define SKH_EINVAL -1    /* invalid argument (NULL pointer, bad enum, capacity too small) */
#define SKH_EHIP -2      /* HIP runtime error */
#define SKH_ENOMEM -3    /* device allocation failed */
#define SKH_ENODEV -4    /* no usable gfx950 device */

#define SKH_DESC_LEN 128 /* features.rs:494-495 DESC_LEN */

/* Distance precision of ratio_match. Every precision returns the reference's
 * match list bit-exactly: the low-precision forms rank with MFMA and re-score
 * every decision that their rigorous error bound cannot settle with the
 * reference's FP32 formula. */
#define SKH_PREC_FP32 0  /* exact FP32 VALU kernel, the reference's sq_dist order */
#define SKH_PREC_BF16 1  /* BF16 MFMA ranking + FP32 re-score */
#define SKH_PREC_FP8 2   /* FP8 e4m3 (per-descriptor scale) MFMA ranking + FP32 re-score */

/* 1 if this build implements `precision` (one of SKH_PREC_*), else 0. */
int skh_precision_supported(int precision);

typedef struct skh_ctx skh_ctx;

/* Creates a context on HIP device `device` (a stream, grow-only scratch).
 * Thread safety: calls on one context are serialised by an internal mutex;
 * use one context per thread for concurrency. */
int skh_ctx_create(int device, skh_ctx** out);
int skh_ctx_destroy(skh_ctx* ctx);
/* MFMA ranking tile (rows per workgroup x columns per chunk). SKH_TILE_AUTO
 * lets the context pick per pair: from its tuned table (skh_ctx_autotune or
 * skh_ctx_set_tile_table) by the pair's size class, else 256x64 (128x64 for
 * at most 128 rows). A fixed tile applies to every pair. */
#define SKH_TILE_AUTO -1
#define SKH_TILE_128x64 0
#define SKH_TILE_128x128 1
#define SKH_TILE_256x64 2
#define SKH_TILE_256x128 3
int skh_ctx_set_tile(skh_ctx* ctx, int tile);
/* Tile autotuner (the batched driver's tile choice, benches/pipeline.rs:219-305
 * runs every pair the same way on the CPU). Pairs fall into 25 size classes:
 * rows and columns each in one of 5 buckets (<= 192, <= 384, <= 1024,
 * <= 3072, more), class = 5 * bucket(rows) + bucket(columns). For each class
 * the tuner times the four tiles' ranking kernel on a batch of synthetic
 * pairs of that class (a few seconds in all), stores the fastest in the
 * context's table for `precision` (SKH_PREC_BF16 or SKH_PREC_FP8) and copies
 * it to table (25 entries, SKH_TILE_*) and the times to ms (25 x 4, ms per
 * launch, class-major); either may be NULL. Results never depend on the tile. */
int skh_ctx_autotune(skh_ctx* ctx, int precision, int32_t* table, double* ms);
/* Loads a table (25 SKH_TILE_* entries) for `precision`, e.g. one saved from
 * an earlier skh_ctx_autotune; NULL clears it. */
int skh_ctx_set_tile_table(skh_ctx* ctx, int precision, const int32_t* table);
/* The tile the context picks for each of the 25 classes; *tuned = 1 if a
 * table is loaded (else the defaults are reported). */
int skh_ctx_tile_table(skh_ctx* ctx, int precision, int32_t* table, int* tuned);
/* The HIP stream (hipStream_t) the context launches on. */
void* skh_ctx_stream(skh_ctx* ctx);
/* Last error message of this context (static storage owned by ctx). */
const char* skh_ctx_last_error(skh_ctx* ctx);

/* candidate_pairs (matching.rs:18-40). views: n (cam, pos) pairs, 2n uint64.
 * Writes up to `cap` pairs (2 uint64 each) to out_ij, ascending (i, then j),
 * and the total count to *out_len. SKH_EINVAL if cap < total (out_len still
 * holds the total so the caller can size the buffer). Host-only. */
int skh_candidate_pairs(const uint64_t* views, size_t n, uint64_t temporal, uint64_t cross,
                        uint64_t pow2_max, uint64_t* out_ij, size_t cap, size_t* out_len);

/* PairSchedule (matching.rs:73-99) with its CrossSchedule (:47-71):
 * cross_kind SKH_CROSS_SPEC pairs different cameras at position differences
 * 0..=range; SKH_CROSS_FORMATION pairs the front camera (0) at position p
 * with the right (1) at p + right_min ..= p + max and the left (2) at
 * p + left_min ..= p + max, every step-th position. */
#define SKH_CROSS_SPEC 0
#define SKH_CROSS_FORMATION 1
typedef struct {
    uint64_t temporal, pow2_max;
    int32_t cross_kind;
    int32_t pad;
    uint64_t range;                          /* SKH_CROSS_SPEC */
    uint64_t right_min, left_min, max, step; /* SKH_CROSS_FORMATION */
} skh_pair_schedule;
/* PairSchedule::default() (matching.rs:81-89): temporal 5, pow2_max 16,
 * CrossSchedule::FORMATION (20, 20, 40, 4). */
void skh_pair_schedule_default(skh_pair_schedule* sch);
/* PairSchedule::spec() (matching.rs:91-98): cross Spec { range: 4 }. */
void skh_pair_schedule_spec(skh_pair_schedule* sch);
/* scheduled_pairs (matching.rs:103-142), the production pair rule
 * (pipeline.rs:486, sparse.rs:311): views as for skh_candidate_pairs, pairs
 * (i, j) with i < j, sorted, no duplicates; same capacity contract. */
int skh_scheduled_pairs(const uint64_t* views, size_t n, const skh_pair_schedule* sch,
                        uint64_t* out_ij, size_t cap, size_t* out_len);

/* verified_matches (tracks.rs:56-69): the (i, j) of `matches` (2 * n_matches
 * uint32) whose inlier flag is set, in order, into out_ij; *out_len = their
 * count. A length mismatch (n_matches != n_inliers) is the reference's
 * Err((n_matches, n_inliers)): SKH_EINVAL, *out_len untouched. Host-only. */
int skh_verified_matches(const uint32_t* matches, size_t n_matches, const uint8_t* inliers,
                         size_t n_inliers, uint32_t* out_ij, size_t* out_len);

/* ---- multi-view tracks (tracks.rs:127-800) ---- */
/* ConflictPolicy (tracks.rs:103-107). */
#define SKH_TRACKS_DROP 0
#define SKH_TRACKS_SPLIT 1
/* TrackConfig (tracks.rs:109-125). */
typedef struct {
    int32_t policy; /* SKH_TRACKS_DROP or SKH_TRACKS_SPLIT */
    int32_t pad;
    uint64_t min_length, max_tracks;
} skh_track_cfg;
/* TrackStats (tracks.rs:145-155). */
typedef struct {
    uint64_t edges, invalid_matches, conflicts, vote_cuts, too_short, inconsistent, truncated,
        tracks;
} skh_track_stats;
/* TrackConfig::default() (tracks.rs:118-125): Split, min_length 3,
 * max_tracks MAX_TRACKS = 100000. */
void skh_track_cfg_default(skh_track_cfg* cfg);
/* build_tracks (tracks.rs:190-508); replaces the Rust call
 * build_tracks(&pairs, &keypoints, &cfg).
 * Pair p (PairMatches): images image_a[p], image_b[p]; matches (fa, fb) =
 * m_ab[2 k], m_ab[2 k + 1] for k in [m_off[p], m_off[p + 1]). Image i's
 * keypoints (x, y) = kp[2 x], kp[2 x + 1] for x in [kp_off[i], kp_off[i + 1]).
 * Tracks out as CSR: track t's observations (image, feature) = obs[2 q],
 * obs[2 q + 1] for q in [track_off[t], track_off[t + 1]) (the pixel is
 * keypoints[image][feature]). *n_tracks and *stats are always written; if
 * track_cap (entries of track_off, n_tracks + 1 needed) or obs_cap
 * (observations) is short the call returns SKH_EINVAL (total keypoints + 1
 * and total keypoints always suffice). The Split policy's displacement
 * filter runs on the device; blocks until done. */
int skh_build_tracks(skh_ctx* ctx, size_t n_pairs, const uint64_t* image_a,
                     const uint64_t* image_b, const uint64_t* m_off, const uint64_t* m_ab,
                     size_t n_images, const uint64_t* kp_off, const double* kp,
                     const skh_track_cfg* cfg, uint64_t* track_off, size_t track_cap,
                     uint64_t* obs, size_t obs_cap, size_t* n_tracks, skh_track_stats* stats);

/* ratio_match (matching.rs:167-215) on host buffers.
 * a: na x 128 f32 row-major, b: nb x 128 f32 (Feature.desc gathered
 * contiguously). out_ij: 2*cap uint32 (i, j), ascending in i; *out_len = count.
 * cap >= na always suffices (min(na, nb) when mutual). Blocks until the result
 * is on the host. */
int skh_ratio_match(skh_ctx* ctx, const float* a, size_t na, const float* b, size_t nb,
                    float ratio, int mutual, int precision, uint32_t* out_ij, size_t cap,
                    size_t* out_len);

/* ratio_match on device buffers (inputs already resident in HBM). d_out_ij
 * must hold 2*na uint32; d_out_len one uint32. Asynchronous on the context
 * stream; the caller synchronises (skh_ctx_stream). */
int skh_ratio_match_dev(skh_ctx* ctx, const float* d_a, size_t na, const float* d_b, size_t nb,
                        float ratio, int mutual, int precision, uint32_t* d_out_ij,
                        uint32_t* d_out_len);

/* skh_ratio_match_dev without allocating: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, results and asynchrony as skh_ratio_match_dev; its working
 * buffers are one fixed device array of about 9.2 MB per GPU, held until the
 * process ends, so na and nb are limited to 8192 rows each (the reference's
 * default feature cap, features.rs:285); larger inputs return SKH_EINVAL.
 * Every call on a GPU shares that array and leaves its contents behind:
 * calls that overlap on one GPU (several contexts or streams at once) race. */
int skh_ratio_match_dev_noalloc(skh_ctx* ctx, const float* d_a, size_t na, const float* d_b,
                                size_t nb, float ratio, int mutual, int precision,
                                uint32_t* d_out_ij, uint32_t* d_out_len);

/* Exact per-row state of ratio_match for diagnostics and tests: row_j1
 * (UINT32_MAX = none), row_d1, row_d2 (na each) and the column winner
 * col_i (UINT32_MAX = none), col_d (nb each), all host buffers, computed
 * with the reference's FP32 formula. */
int skh_ratio_match_state(skh_ctx* ctx, const float* a, size_t na, const float* b, size_t nb,
                          uint32_t* row_j1, float* row_d1, float* row_d2, uint32_t* col_i,
                          float* col_d);

/* RansacConfig (matching.rs:762-793). */
typedef struct {
    double threshold_px;     /* inlier Sampson distance threshold (px, not squared) */
    double min_inlier_ratio; /* below this inlier ratio -> None */
    uint64_t max_iters;
    double confidence;       /* adaptive termination confidence */
    uint64_t seed;
    uint64_t min_inliers;
    double max_chance_prob;  /* <= 0 disables the significance gate */
} skh_ransac_cfg;

/* ModelSelection (matching.rs:380-394) plus a presence flag (the reference's
 * Option). model: 0 = Fundamental, 1 = Homography. */
typedef struct {
    int32_t has_selection;
    int32_t model;
    double gric_f, gric_h;
    uint64_t parallax, parallax_needed;
    double sigma_px;
} skh_model_sel;

/* RansacConfig::default() (matching.rs:781-793): 1.5 px, 0.2, 2000, 0.999,
 * seed 1, 15 inliers, 1e-6. */
void skh_ransac_cfg_default(skh_ransac_cfg* cfg);

/* ransac_fundamental (matching.rs:941-1104) on host buffers.
 * x1, x2: n x 2 f64 (Vector2<f64> slices, nalgebra column layout [x, y]).
 * On Some: *found = 1, F_rowmajor = F (unit Frobenius norm, rank 2; the Rust
 * wrapper builds Matrix3::new from the row-major values), inliers[k] = 1 for
 * inliers (n bytes), *sel = the model selection (has_selection = 0 for the
 * reference's selection: None). On None: *found = 0. Blocks until done. */
int skh_ransac_fundamental(skh_ctx* ctx, const double* x1, const double* x2, size_t n,
                           const skh_ransac_cfg* cfg, double* F_rowmajor, uint8_t* inliers,
                           skh_model_sel* sel, int* found);

/* Hypotheses drawn by the context's last skh_ransac_fundamental call
 * (diagnostics; the reference's adaptive stopping point). */
uint64_t skh_ctx_last_ransac_iters(skh_ctx* ctx);

/* ---- batched per-pair driver (benches/pipeline.rs:219-305) ----
 * All pairs of a scene: ratio_match, the Feature -> pixel hand-off and
 * ransac_fundamental, as one captured hipGraph on the context stream. The
 * plan fixes the views, the pair list and the parameters; every run
 * recomputes everything from the device-resident descriptors and keypoints.
 *
 * views: n_views feature ranges [view_offset[v], view_offset[v+1]) into
 *   d_desc (total x 128 f32) and d_kp (total x 2 f32, Keypoint.x/.y), both
 *   device pointers that must stay valid for the plan's lifetime.
 * pairs: n_pairs (i, j) view indices (host), e.g. skh_candidate_pairs output.
 * Each pair runs ratio_match(view i, view j, ratio, mutual) and, on its match
 * list, ransac_fundamental with x1 = kp_i + px_offset, x2 = kp_j + px_offset
 * (benches/pipeline.rs:276-283 uses 0.0; README.md:376-379 uses 0.5). */
typedef struct {
    uint32_t n_views;
    const uint64_t* view_offset; /* host, n_views + 1 */
    const float* d_desc;         /* device */
    const float* d_kp;           /* device */
    uint32_t n_pairs;
    const uint32_t* pairs;       /* host, 2 * n_pairs */
    float ratio;
    int32_t mutual;
    int32_t precision;           /* SKH_PREC_* */
    int32_t run_ransac;          /* 0: matching only */
    double px_offset;
    skh_ransac_cfg ransac;
} skh_batch_desc;

typedef struct skh_plan skh_plan;

int skh_plan_create(skh_ctx* ctx, const skh_batch_desc* desc, skh_plan** out);
int skh_plan_destroy(skh_plan* plan);
/* Enqueues one run (graph launch) on the context stream; asynchronous. */
int skh_plan_run(skh_plan* plan);
/* Waits for the plan's runs to finish. */
int skh_plan_sync(skh_plan* plan);
/* Per-pair match capacity offsets: match_base[p] (host, n_pairs + 1). The
 * matches of pair p live at [match_base[p], match_base[p] + match_len[p]). */
int skh_plan_match_base(skh_plan* plan, uint64_t* match_base);
/* Copies results to host buffers (any may be NULL): match_len (n_pairs),
 * match_ij (2 * match_base[n_pairs] uint32), inliers (match_base[n_pairs]
 * bytes, aligned with match_ij), F (9 * n_pairs, row-major), found (n_pairs:
 * 1 = Some, 0 = None), sel (n_pairs). Synchronises first. Returns SKH_EHIP
 * (after copying) if any pair's RANSAC found its replay inconsistent, an
 * internal error that the single-pair call reports the same way. */
int skh_plan_results(skh_plan* plan, uint32_t* match_len, uint32_t* match_ij, uint8_t* inliers,
                     double* F, int32_t* found, skh_model_sel* sel);
/* Device time (ms) of the named stage in the plan's last run, measured with
 * HIP events on the context stream when timing is enabled: stage 0 = whole
 * run, 1 = matching, 2 = RANSAC. Returns -1 if unavailable. */
double skh_plan_stage_ms(skh_plan* plan, int stage);
/* Average duration (ms) of the distance kernel alone (MFMA ranking, or the
 * exact FP32 kernel), launched `reps` times back to back between HIP events
 * on the context stream, on the inputs of the plan's last run (call after a
 * run). Synchronises. Returns -1 on error. */
double skh_plan_time_distance(skh_plan* plan, int reps);
/* Approximate ranking state of the last run of a BF16 or FP8 plan, for error
 * checks and diagnostics (SKH_EINVAL for an FP32 plan). The rows of pair p
 * occupy [R_p, R_p + na_p), R_p = the sum of na over the pairs before p (the
 * columns likewise with nb). row_k1 / row_k2: the row's two smallest keys
 * (key ~ d^2 / 2 with the column-tile id in its lowest mantissa bits);
 * row_j1: the column of the smallest; col_top2: (c1, c2) per column, the
 * column's two smallest keys; eps[p]: the pair's error bound, |2 key - d^2| <=
 * eps for every element of the pair, d^2 the exact squared distance (infinity
 * when the pair is scored exactly); *queued: rows the run sent to the exact
 * FP32 re-score. Any pointer may be NULL. Synchronises. */
int skh_plan_lowp_state(skh_plan* plan, float* row_k1, uint32_t* row_j1, float* row_k2,
                        float* col_top2, double* eps, uint64_t* queued);
/* Verified matches of every pair of the last run, the consumer's input
 * (tracks.rs:56-69 applied per pair; the pipeline keeps a pair only when
 * RANSAC found a model): verified_len[p] = the inlier count when found[p] = 1,
 * else 0; verified_ij holds pair p's verified (i, j) at
 * [match_base[p], match_base[p] + verified_len[p]) (2 uint32 each), in match
 * order. Computed on the device by the RANSAC graph; either pointer may be
 * NULL. Synchronises. */
int skh_plan_verified(skh_plan* plan, uint32_t* verified_len, uint32_t* verified_ij);
/* Pairs of the plan ranked with each tile (4 entries, SKH_TILE_* order; all
 * zero for an FP32 plan). */
int skh_plan_tiles(skh_plan* plan, uint32_t* pairs_per_tile);
/* 1 = record stage events in each run (costs nothing measurable), 0 = off. */
int skh_plan_set_timing(skh_plan* plan, int on);
/* RANSAC diagnostics of the last run (any pointer may be NULL): iters[p] =
 * hypotheses drawn; phase[11 * p + k] = device wall-clock ticks of phase k
 * (0 sampling, 1 hypotheses, 2 local optimisation, 3 replay, 4 pool LM,
 * 5 final refits and gates, 6 line test, 7 model selection, 8 batches,
 * 9 LO lanes, 10 total); *ticks_per_ms = the wall clock's rate. */
int skh_plan_ransac_profile(skh_plan* plan, uint64_t* iters, uint64_t* phase, double* ticks_per_ms);

/* ---- feature detector: detect_and_describe (features.rs:594-648) ----------
 * The producer of the path's descriptors: DoG extrema with sub-pixel /
 * sub-scale refinement, edge test, dominant orientations and the 128-d
 * descriptor, on a single-channel f32 image (GrayImage, row-major, w x h,
 * features.rs:47-52; the Rust wrapper passes GrayImage::data()).
 * Results equal the reference's algorithm in f32 with its operation order;
 * the per-pixel f32 transcendentals (exp, atan2, sin, cos, pow) are the
 * correctly rounded values (the reference's come from the platform libm,
 * which misrounds some of them in the last bit). */
typedef struct {
    uint64_t octaves, scales; /* DetectorConfig (features.rs:259-276), usize as uint64 */
    float sigma0, contrast, edge_ratio;
    uint32_t upscale_fill; /* bool */
    uint64_t max_features, upscale_below;
} skh_detector_cfg;
/* Keypoint (features.rs:295-304). */
typedef struct {
    float x, y, sigma, response, angle;
} skh_keypoint;
/* DetectorConfig::default() (features.rs:277-289): 4 octaves, 3 scales,
 * sigma0 1.6, contrast 0.01, edge ratio 10, 8192 features, upscale below 800. */
void skh_detector_cfg_default(skh_detector_cfg* cfg);
/* detect_and_describe (features.rs:594-648) on a host image: writes the
 * features in the reference's order, keypoints to kps (cap entries) and
 * descriptors to desc (cap x 128 f32; NULL skips them: detect(),
 * features.rs:487-493), and their count to *out_len. cap >= max_features
 * always suffices; SKH_EINVAL if cap is short (*out_len holds the count).
 * This build rejects scales outside 1..14 and Gaussian radii above 64
 * (sigma0 up to ~21). Blocks until the results are on the host. */
int skh_detect_and_describe(skh_ctx* ctx, const float* img, size_t w, size_t h,
                            const skh_detector_cfg* cfg, skh_keypoint* kps, float* desc,
                            size_t cap, size_t* out_len);
/* The same on a device-resident image, writing device outputs (d_desc may be
 * NULL). Returns once the count is known (*out_len on the host); the
 * keypoint and descriptor kernels may still run on the context stream. */
int skh_detect_and_describe_dev(skh_ctx* ctx, const float* d_img, size_t w, size_t h,
                                const skh_detector_cfg* cfg, skh_keypoint* d_kps, float* d_desc,
                                size_t cap, size_t* out_len);

#ifdef __cplusplus
}
#endif
#endif
