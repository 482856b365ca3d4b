// Internal declarations shared by the HIP translation units of libskylens_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace skh {

constexpr int kDescLen = 128;

// Which low-precision ranking kernels this build contains.
constexpr bool kHaveBf16 = true;
constexpr bool kHaveFp8 = true;

// One ratio_match problem: a (na rows) against b (nb rows), both row-major
// [n][128] f32 in device memory. row_base/col_base index the per-row and
// per-column state arrays shared by all pairs of a batch.
struct PairTask {
    const float* a;
    const float* b;
    uint32_t na, nb;
    uint32_t row_base;
    uint32_t col_base;
};

// One workgroup's share of the exact kernel: 64 rows of one pair.
struct RowBlock {
    uint32_t pair;
    uint32_t row0;
};

constexpr uint32_t kNone = 0xffffffffu;
constexpr unsigned long long kColEmpty = 0xffffffffffffffffull;

// Exact FP32 row/column state of ratio_match (matching.rs:167-203).
// row_j1/row_d1/row_d2 are indexed row_base + i; col_key is indexed
// col_base + j and must be preset to kColEmpty. col_key packs
// (float bits of d) << 32 | i, so an unsigned min keeps the reference's
// "smallest d, lowest i on ties" winner.
void launch_exact_state(const PairTask* d_tasks, const RowBlock* d_blocks, uint32_t n_blocks,
                        uint32_t* row_j1, float* row_d1, float* row_d2,
                        unsigned long long* col_key, hipStream_t s);

// Resets n column keys to kColEmpty (a fill kernel, safe under graph capture).
void launch_fill_u64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s);

void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);

// Final decision of matching.rs:204-214 from exact state: writes
// row_match[row_base + i] = j1 when row i is accepted, else kNone.
void launch_exact_decide(const PairTask* d_tasks, uint32_t n_pairs, const uint32_t* row_j1,
                         const float* row_d1, const float* row_d2,
                         const unsigned long long* col_key, float ratio, int mutual,
                         uint32_t* row_match, hipStream_t s);

// Compacts accepted rows of every pair into (i, j) uint32 pairs at
// out_ij + 2 * (out_base[p]) in ascending i, and the count into out_len[p].
void launch_compact(const PairTask* d_tasks, uint32_t n_pairs, const uint32_t* row_match,
                    const uint64_t* out_base, uint32_t* out_ij, uint32_t* out_len,
                    hipStream_t s);

}  // namespace skh
