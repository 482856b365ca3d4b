// Shared declarations of the low-precision (MFMA-ranked) ratio_match.
#pragma once

#include "skh_internal.hpp"

namespace skh {

// Per-view inputs of the error bound (all upper bounds, f32): max |a|,
// max |a - q(a)|, max |q(a)|, max |a|^2, where q(a) is the quantised row in
// real units (bf16, or e4m3 times its scale); bad != 0 if any component is
// non-finite or too large for the bound (such views are scored exactly).
struct ViewStats {
    float max_norm, max_res, max_qnorm, max_sq;
    uint32_t bad, pad;
};

// FP8 row metadata: x = E8M0 code of the row's scale 2^e (e + 127) | E8M0
// code of the norm's fold scale << 8; y = e4m3 bytes hi | mid << 8 | lo << 16
// of |a|^2 / 2^f (hi + mid + lo carry |a|^2 to 2^-11 relative).
struct Fp8Meta {
    uint32_t codes;
    uint32_t fold;
};

struct LowpTask {
    const float* a;            // f32 rows (exact re-scoring)
    const float* b;
    const uint16_t* qa;        // bf16 rows [n][128]
    const uint16_t* qb;
    const uint32_t* nrm_a;     // packed bf16 (hi | lo << 16) of |row|^2
    const uint32_t* nrm_b;
    const uint8_t* q8a;        // e4m3 rows [n][128] (FP8 ranking)
    const uint8_t* q8b;
    const Fp8Meta* meta_a;
    const Fp8Meta* meta_b;
    const ViewStats* sa;
    const ViewStats* sb;
    uint32_t na, nb;
    uint32_t row_base;         // per-row arrays
    uint32_t col_base;         // per-column arrays (col_top2)
    uint64_t part_base;        // column partials: ceil(na / BM) * nb float2
    uint32_t fp8;              // 1: e4m3 ranking (error bound of the FP8 form)
    uint32_t bm;               // rows per row block of this pair's tile
};

struct LowpOut {
    float* row_k1;             // packed key of the row minimum (~ D/2)
    uint32_t* row_j1;          // its column
    float* row_k2;             // second smallest key
    float2* col_part;          // per workgroup column (c1, c2)
    float2* col_top2;          // merged column (c1, c2)
    uint32_t* row_match;       // final j or kNone
    // Exact re-score queue, one contiguous run of entries per pair. Entry
    // (pair, row, j1 | kNone, w): j1 = kNone needs the row scan, else only the
    // column check of j1; w = 0 (open), or, set by the FP8 path's BF16
    // re-rank, 1 (settled: rejected) or 2 (settled: accepted with j1).
    uint4* queue;
    uint32_t* queue_len;       // [0] queue entries, [1] chunks
    uint4* pre;                // per entry: exact (j1, d1 bits, d2 bits, 1) of a gathered row scan
    // Chunks (first entry, entries <= kChunkEntries) of the runs of pairs with
    // at least kGatherMinRows queued rows, for the gathered row scans. Null:
    // no gathered pass (pre unused); the BF16 queue is too sparse to gain.
    uint2* chunks;
};

constexpr uint32_t kChunkEntries = 64;
constexpr uint32_t kGatherMinRows = 16;
// Chunk list capacity of a batch with `rows` a-side rows over `pairs` pairs.
constexpr uint64_t chunk_capacity(uint64_t rows, uint64_t pairs) { return rows / kChunkEntries + pairs; }

enum TileShape : int {
    kTile128x64 = 0,   // 4 waves x 32 rows, 64-column chunks
    kTile128x128 = 1,  // 4 waves x 32 rows, 128-column chunks
    kTile256x64 = 2,   // 4 waves x 64 rows, 64-column chunks
    kTile256x128 = 3,  // 4 waves x 64 rows, 128-column chunks
};
constexpr int kNumTiles = 4;
inline uint32_t tile_rb(int tile) { return (tile == kTile256x64 || tile == kTile256x128) ? 2u : 1u; }
inline uint32_t tile_bn(int tile) { return (tile == kTile128x128 || tile == kTile256x128) ? 128u : 64u; }

uint32_t lowp_block_rows(int tile);
void launch_stats_reset(ViewStats* st, uint32_t n, hipStream_t s);
void launch_quantize(const float* x, uint32_t n, uint16_t* q, uint32_t* nrm, ViewStats* st,
                     hipStream_t s);
void launch_quantize_views(const float* x, const uint64_t* d_off, uint32_t n_views,
                           uint32_t max_rows, uint16_t* q, uint32_t* nrm, ViewStats* st,
                           hipStream_t s);
void launch_quantize_fp8(const float* x, uint32_t n, uint8_t* q, Fp8Meta* meta, ViewStats* st,
                        hipStream_t s);
void launch_quantize_views_fp8(const float* x, const uint64_t* d_off, uint32_t n_views,
                               uint32_t max_rows, uint8_t* q, Fp8Meta* meta, ViewStats* st,
                               hipStream_t s);
// MFMA ranking of every row block: BF16 (fp8 = false) or FP8 e4m3 operands.
void launch_mfma_rank(int tile, bool fp8, const LowpTask* tasks, const RowBlock* blocks,
                      uint32_t n_blocks, const LowpOut& out, uint32_t tbits, hipStream_t s);
// Per-pair error bound eps (>= |2 key - D_ref| over every element of the pair).
void launch_pair_eps(const LowpTask* tasks, uint32_t n_pairs, uint32_t tbits, double* eps,
                     hipStream_t s);
void launch_colmerge(const LowpTask* tasks, uint32_t n_pairs, const LowpOut& out, hipStream_t s);

// Size classes of the tile table: rows and columns of a pair each fall in one
// of kSizeBuckets buckets (<= 192, <= 384, <= 1024, <= 3072, more).
constexpr int kSizeBuckets = 5;
constexpr int kTileClasses = kSizeBuckets * kSizeBuckets;
inline int size_bucket(uint32_t n) {
    return n <= 192 ? 0 : n <= 384 ? 1 : n <= 1024 ? 2 : n <= 3072 ? 3 : 4;
}
inline int tile_class(uint32_t na, uint32_t nb) { return kSizeBuckets * size_bucket(na) + size_bucket(nb); }
void launch_decide(const LowpTask* tasks, uint32_t n_pairs, const LowpOut& out, float ratio,
                   int mutual, uint32_t tbits, hipStream_t s);
// FP8 plans: BF16 re-rank of the row decisions the FP8 bound left open (per
// queue chunk), settling entries with the BF16 bound before the exact passes.
// tasks_b: the pairs' BF16 tasks (bf16 rows, norms, statistics); tbits_g: tile
// bits of 64-column chunks over the widest view.
void launch_rerank_bf16(const LowpTask* tasks_b, const LowpTask* tasks_8, const LowpOut& out,
                        float ratio, int mutual, uint32_t tbits_8, uint32_t tbits_g,
                        uint32_t grid, hipStream_t s);
void launch_rescore(const LowpTask* tasks, const LowpOut& out, float ratio, int mutual,
                    uint32_t tbits, uint32_t grid, hipStream_t s);

// Bits needed for a column-tile id of a pair with up to nb_max columns swept in
// chunks of bn columns (tiles of 32): the last chunk's padding tiles count too,
// so no tile id wraps onto a real column.
inline uint32_t tile_bits(uint64_t nb_max, uint32_t bn) {
    uint64_t tiles = (nb_max + bn - 1) / bn * (bn / 32);
    uint32_t b = 1;
    while ((1ull << b) < tiles) ++b;
    return b;
}

}  // namespace skh
