// Host orchestration of the low-precision (BF16 or FP8) single-pair ratio_match.
#include <algorithm>
#include <vector>

#include "skh_ctx.hpp"
#include "skh_lowp.hpp"
#include "skylens_hip.h"

namespace skh {

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
        if ((rc = ensure(out_base, sizeof(uint32_t), "out_base"))) return rc;
        launch_fill_u32(d_out_len, 1, 0u, stream);
        return hip(hipGetLastError(), "fill");
    }
    if (nb > kLowpMaxCols)
        return ratio_match_dev(d_a, na, d_b, nb, ratio, mutual, SKH_PREC_FP32, d_out_ij, d_out_len);

    const int tile = tile_override >= 0 ? tile_override : (na <= 128 ? kTile128x64 : kTile256x64);
    const uint32_t bm = lowp_block_rows(tile);
    const uint32_t nblk = (uint32_t)((na + bm - 1) / bm);
    const uint32_t tbits = tile_bits(nb, tile_bn(tile));
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
        (rc = ensure(queue, sizeof(uint4) * na, "queue")) ||
        (rc = ensure(queue_len, 4, "queue_len")) || (rc = ensure(out_base, 8, "out_base")) ||
        (rc = ensure(tasks, sizeof(PairTask), "tasks")))
        return rc;

    ViewStats* st = stats.as<ViewStats>();
    LowpTask lt{};
    lt.a = d_a;
    lt.b = d_b;
    if (fp8) {
        lt.q8a = q8_a.as<uint8_t>();
        lt.q8b = q8_b.as<uint8_t>();
        lt.meta_a = meta_a.as<Fp8Meta>();
        lt.meta_b = meta_b.as<Fp8Meta>();
        lt.fp8 = 1;
    } else {
        lt.qa = q_a.as<uint16_t>();
        lt.qb = q_b.as<uint16_t>();
        lt.nrm_a = nrm_a.as<uint32_t>();
        lt.nrm_b = nrm_b.as<uint32_t>();
    }
    lt.sa = st;
    lt.sb = st + 1;
    lt.na = (uint32_t)na;
    lt.nb = (uint32_t)nb;
    PairTask pt{d_a, d_b, (uint32_t)na, (uint32_t)nb, 0u, 0u};
    h_blocks.resize(nblk);
    for (uint32_t k = 0; k < nblk; ++k) h_blocks[k] = RowBlock{0u, k * bm};
    if ((rc = hip(hipMemcpyAsync(ltasks.ptr, &lt, sizeof lt, hipMemcpyHostToDevice, stream),
                  "copy ltask")) ||
        (rc = hip(hipMemcpyAsync(tasks.ptr, &pt, sizeof pt, hipMemcpyHostToDevice, stream),
                  "copy task")) ||
        (rc = hip(hipMemcpyAsync(blocks.ptr, h_blocks.data(), sizeof(RowBlock) * nblk,
                                 hipMemcpyHostToDevice, stream),
                  "copy blocks")))
        return rc;
    // The host arrays above must outlive the async copies.
    if ((rc = hip(hipStreamSynchronize(stream), "sync"))) return rc;

    LowpOut o{};
    o.row_k1 = row_k1.as<float>();
    o.row_j1 = row_j1.as<uint32_t>();
    o.row_k2 = row_k2.as<float>();
    o.col_part = col_part.as<float2>();
    o.col_top2 = col_top2.as<float2>();
    o.row_match = row_match.as<uint32_t>();
    o.queue = queue.as<uint4>();
    o.queue_len = queue_len.as<uint32_t>();

    launch_stats_reset(st, 2, stream);
    launch_fill_u32(o.queue_len, 1, 0u, stream);
    launch_fill_u64(reinterpret_cast<unsigned long long*>(out_base.ptr), 1, 0ull, stream);
    if (fp8) {
        launch_quantize_fp8(d_a, (uint32_t)na, q8_a.as<uint8_t>(), meta_a.as<Fp8Meta>(), st, stream);
        launch_quantize_fp8(d_b, (uint32_t)nb, q8_b.as<uint8_t>(), meta_b.as<Fp8Meta>(), st + 1,
                            stream);
    } else {
        launch_quantize(d_a, (uint32_t)na, q_a.as<uint16_t>(), nrm_a.as<uint32_t>(), st, stream);
        launch_quantize(d_b, (uint32_t)nb, q_b.as<uint16_t>(), nrm_b.as<uint32_t>(), st + 1, stream);
    }
    launch_mfma_rank(tile, fp8, ltasks.as<LowpTask>(), blocks.as<RowBlock>(), nblk, o, tbits,
                     stream);
    launch_colmerge(ltasks.as<LowpTask>(), 1, o, bm, stream);
    launch_decide(ltasks.as<LowpTask>(), 1, o, ratio, mutual, tbits, stream);
    launch_rescore(ltasks.as<LowpTask>(), o, ratio, mutual, tbits, (uint32_t)std::max(64, 4 * num_cu),
                   stream);
    launch_compact(tasks.as<PairTask>(), 1, o.row_match, out_base.as<uint64_t>(), d_out_ij,
                   d_out_len, stream);
    return hip(hipGetLastError(), "launch lowp");
}

}  // namespace skh
