// Host orchestration of the single-pair ratio_match passes: the exact FP32
// pass and the low-precision (BF16 or FP8) pass, on a set of device buffers.
#include <algorithm>

#include "skh_ctx.hpp"
#include "skh_lowp.hpp"
#include "skh_single.hpp"
#include "skylens_hip.h"

namespace skh {
namespace {

// Writes the single pair's task records and its row-block list on the device,
// so a call needs no host-to-device copy (and no host buffer to keep alive).
__global__ void single_setup_kernel(LowpTask lt, PairTask pt, uint32_t nblk, uint32_t bm,
                                    LowpTask* __restrict__ d_lt, PairTask* __restrict__ d_pt,
                                    RowBlock* __restrict__ d_blocks) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        if (d_lt) *d_lt = lt;
        *d_pt = pt;
    }
    if (k < nblk) d_blocks[k] = RowBlock{0u, k * bm};
}

}  // namespace

void launch_single_setup(const LowpTask* lt, const PairTask& pt, uint32_t nblk, uint32_t bm,
                         LowpTask* d_lt, PairTask* d_pt, RowBlock* d_blocks, hipStream_t s) {
    const uint32_t n = std::max(nblk, 1u);
    hipLaunchKernelGGL(single_setup_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       lt ? *lt : LowpTask{}, pt, nblk, bm, lt ? d_lt : nullptr, d_pt, d_blocks);
}

int Ctx::pick_tile(int precision, size_t na, size_t nb) const {
    if (tile_override >= 0) return tile_override;
    const int p = precision == SKH_PREC_FP8 ? 1 : 0;
    if (tile_tuned[p]) return tile_table[p][tile_class((uint32_t)na, (uint32_t)nb)];
    return na <= 128 ? kTile128x64 : kTile256x64;
}

// The whole low-precision single-pair pass on the buffers of `B`: set-up,
// quantisation, MFMA ranking, column merge, certain decisions, exact
// re-score, compaction into d_out_ij / d_out_len. Asynchronous on `s`; no
// allocation and no host synchronisation. Needs na >= 1 and nb >= 2.
void enqueue_lowp_single(const SingleBufs& B, const float* d_a, uint32_t na, const float* d_b,
                         uint32_t nb, float ratio, int mutual, bool fp8, int tile, int num_cu,
                         uint32_t* d_out_ij, uint32_t* d_out_len, hipStream_t s) {
    const uint32_t bm = lowp_block_rows(tile);
    const uint32_t nblk = (na + bm - 1) / bm;
    const uint32_t tbits = tile_bits(nb, tile_bn(tile));
    LowpTask lt{};
    lt.a = d_a;
    lt.b = d_b;
    if (fp8) {
        lt.q8a = B.q8_a;
        lt.q8b = B.q8_b;
        lt.meta_a = B.meta_a;
        lt.meta_b = B.meta_b;
        lt.fp8 = 1;
    } else {
        lt.qa = B.q_a;
        lt.qb = B.q_b;
        lt.nrm_a = B.nrm_a;
        lt.nrm_b = B.nrm_b;
    }
    lt.sa = B.stats;
    lt.sb = B.stats + 1;
    lt.na = na;
    lt.nb = nb;
    lt.bm = bm;
    const PairTask pt{d_a, d_b, na, nb, 0u, 0u};
    launch_single_setup(&lt, pt, nblk, bm, B.ltask, B.task, B.blocks, s);

    LowpOut o{};
    o.row_k1 = B.row_k1;
    o.row_j1 = B.row_j1;
    o.row_k2 = B.row_k2;
    o.col_part = B.col_part;
    o.col_top2 = B.col_top2;
    o.row_match = B.row_match;
    o.queue = B.queue;
    o.queue_len = B.queue_len;
    o.pre = B.pre;
    o.chunks = fp8 ? B.chunks : nullptr;

    launch_stats_reset(B.stats, 2, s);
    launch_fill_u32(o.queue_len, 2, 0u, s);
    launch_fill_u64(reinterpret_cast<unsigned long long*>(B.out_base), 1, 0ull, s);
    if (fp8) {
        launch_quantize_fp8(d_a, na, B.q8_a, B.meta_a, B.stats, s);
        launch_quantize_fp8(d_b, nb, B.q8_b, B.meta_b, B.stats + 1, s);
    } else {
        launch_quantize(d_a, na, B.q_a, B.nrm_a, B.stats, s);
        launch_quantize(d_b, nb, B.q_b, B.nrm_b, B.stats + 1, s);
    }
    launch_mfma_rank(tile, fp8, B.ltask, B.blocks, nblk, o, tbits, s);
    launch_colmerge(B.ltask, 1, o, s);
    launch_decide(B.ltask, 1, o, ratio, mutual, tbits, s);
    launch_rescore(B.ltask, o, ratio, mutual, tbits, (uint32_t)std::max(64, 4 * num_cu), s);
    launch_compact(B.task, 1, o.row_match, B.out_base, d_out_ij, d_out_len, s);
}

// The exact FP32 pass on the buffers of `B` (row state, column keys, decision,
// compaction). Asynchronous; no allocation, no host synchronisation.
void enqueue_exact_single(const SingleBufs& B, const float* d_a, uint32_t na, const float* d_b,
                          uint32_t nb, float ratio, int mutual, uint32_t* d_out_ij,
                          uint32_t* d_out_len, hipStream_t s) {
    const uint32_t nblk = (na + 63) / 64;
    const PairTask pt{d_a, d_b, na, nb, 0u, 0u};
    launch_single_setup(nullptr, pt, nblk, 64u, nullptr, B.task, B.blocks, s);
    launch_fill_u64(B.col_key, nb, kColEmpty, s);
    launch_exact_state(B.task, B.blocks, nblk, B.row_j1, B.row_d1, B.row_d2, B.col_key, s);
    launch_fill_u64(reinterpret_cast<unsigned long long*>(B.out_base), 1, 0ull, s);
    launch_exact_decide(B.task, 1, B.row_j1, B.row_d1, B.row_d2, B.col_key, ratio, mutual,
                        B.row_match, s);
    launch_compact(B.task, 1, B.row_match, B.out_base, d_out_ij, d_out_len, s);
}

// Largest b the packed tile id supports without eroding the key precision
// below 2^-11 relative.
constexpr size_t kLowpMaxCols = 131072;

int Ctx::ratio_match_lowp(const float* d_a, size_t na, const float* d_b, size_t nb, float ratio,
                          int mutual, int precision, uint32_t* d_out_ij, uint32_t* d_out_len) {
    int rc;
    if (precision != SKH_PREC_BF16 && precision != SKH_PREC_FP8)
        return fail(SKH_EINVAL, "precision %d not built", precision);
    const bool fp8 = precision == SKH_PREC_FP8;
    if (na == 0 || nb < 2) {
        // No row can pass (matching.rs:206 needs a finite second neighbour).
        launch_fill_u32(d_out_len, 1, 0u, stream);
        return hip(hipGetLastError(), "fill");
    }
    if (nb > kLowpMaxCols)
        return ratio_match_dev(d_a, na, d_b, nb, ratio, mutual, SKH_PREC_FP32, d_out_ij, d_out_len);

    const int tile = pick_tile(precision, na, nb);
    const uint32_t bm = lowp_block_rows(tile);
    const uint32_t nblk = (uint32_t)((na + bm - 1) / bm);
    if (fp8) {
        if ((rc = ensure(q8_a, kDescLen * na, "q8_a")) || (rc = ensure(q8_b, kDescLen * nb, "q8_b")) ||
            (rc = ensure(meta_a, sizeof(Fp8Meta) * na, "meta_a")) ||
            (rc = ensure(meta_b, sizeof(Fp8Meta) * nb, "meta_b")))
            return rc;
    } else if ((rc = ensure(q_a, 2 * kDescLen * na, "q_a")) ||
               (rc = ensure(q_b, 2 * kDescLen * nb, "q_b")) || (rc = ensure(nrm_a, 4 * na, "nrm_a")) ||
               (rc = ensure(nrm_b, 4 * nb, "nrm_b"))) {
        return rc;
    }
    if ((rc = ensure(stats, 2 * sizeof(ViewStats), "stats")) ||
        (rc = ensure(ltasks, sizeof(LowpTask), "ltasks")) ||
        (rc = ensure(blocks, sizeof(RowBlock) * nblk, "blocks")) ||
        (rc = ensure(row_k1, 4 * na, "row_k1")) || (rc = ensure(row_j1, 4 * na, "row_j1")) ||
        (rc = ensure(row_k2, 4 * na, "row_k2")) ||
        (rc = ensure(col_part, sizeof(float2) * (size_t)nblk * nb, "col_part")) ||
        (rc = ensure(col_top2, sizeof(float2) * nb, "col_top2")) ||
        (rc = ensure(row_match, 4 * na, "row_match")) ||
        (rc = ensure(queue, sizeof(uint4) * na, "queue")) || (rc = ensure(pre, sizeof(uint4) * na, "pre")) ||
        (rc = ensure(chunks, sizeof(uint2) * chunk_capacity(na, 1), "chunks")) ||
        (rc = ensure(queue_len, 8, "queue_len")) || (rc = ensure(out_base, 8, "out_base")) ||
        (rc = ensure(tasks, sizeof(PairTask), "tasks")))
        return rc;

    SingleBufs B{};
    B.q_a = q_a.as<uint16_t>();
    B.q_b = q_b.as<uint16_t>();
    B.nrm_a = nrm_a.as<uint32_t>();
    B.nrm_b = nrm_b.as<uint32_t>();
    B.q8_a = q8_a.as<uint8_t>();
    B.q8_b = q8_b.as<uint8_t>();
    B.meta_a = meta_a.as<Fp8Meta>();
    B.meta_b = meta_b.as<Fp8Meta>();
    B.stats = stats.as<ViewStats>();
    B.ltask = ltasks.as<LowpTask>();
    B.task = tasks.as<PairTask>();
    B.blocks = blocks.as<RowBlock>();
    B.row_k1 = row_k1.as<float>();
    B.row_j1 = row_j1.as<uint32_t>();
    B.row_k2 = row_k2.as<float>();
    B.col_part = col_part.as<float2>();
    B.col_top2 = col_top2.as<float2>();
    B.row_match = row_match.as<uint32_t>();
    B.queue = queue.as<uint4>();
    B.queue_len = queue_len.as<uint32_t>();
    B.pre = pre.as<uint4>();
    B.chunks = chunks.as<uint2>();
    B.out_base = out_base.as<uint64_t>();
    enqueue_lowp_single(B, d_a, (uint32_t)na, d_b, (uint32_t)nb, ratio, mutual, fp8, tile, num_cu,
                        d_out_ij, d_out_len, stream);
    return hip(hipGetLastError(), "launch lowp");
}

}  // namespace skh
