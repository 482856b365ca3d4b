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
 *   per-pair driver   benches/pipeline.rs:219-305 -> skh_match_pairs_batched
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
 * lets the context pick (the autotuned table when one has been loaded). */
#define SKH_TILE_AUTO -1
#define SKH_TILE_128x64 0
#define SKH_TILE_128x128 1
#define SKH_TILE_256x64 2
#define SKH_TILE_256x128 3
int skh_ctx_set_tile(skh_ctx* ctx, int tile);
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

/* Exact per-row state of ratio_match for diagnostics and tests: row_j1
 * (UINT32_MAX = none), row_d1, row_d2 (na each) and the column winner
 * col_i (UINT32_MAX = none), col_d (nb each), all host buffers, computed
 * with the reference's FP32 formula. */
int skh_ratio_match_state(skh_ctx* ctx, const float* a, size_t na, const float* b, size_t nb,
                          uint32_t* row_j1, float* row_d1, float* row_d2, uint32_t* col_i,
                          float* col_d);

#ifdef __cplusplus
}
#endif
#endif
