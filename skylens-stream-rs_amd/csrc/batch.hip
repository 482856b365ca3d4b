// Batched per-pair driver: all candidate pairs of a scene in one captured
// hipGraph (the MI355X counterpart of benches/pipeline.rs:219-305, which runs
// ratio_match pair by pair and ransac_fundamental over pairs in parallel).
//
// Scheduling: every pair contributes ceil(na / BM) row-block workgroups to the
// distance kernel. Pairs are assigned to the 8 XCDs longest-first (greedy
// LPT on na * nb), and the launch order interleaves the per-XCD lists so that
// workgroup b lands on XCD b % 8 (the hardware's round-robin dispatch): all
// row blocks of a pair share one XCD's L2, where the pair's b descriptors
// stay resident. RANSAC runs one 64-lane workgroup per pair.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "skh_ctx.hpp"
#include "skh_lowp.hpp"
#include "skh_ransac.hpp"
#include "skylens_hip.h"

namespace skh {
namespace {

constexpr int kXcds = 8;

// Gathers each pair's matched keypoints into the RANSAC input
// (benches/pipeline.rs:276-283: f32 Keypoint coordinates widened to f64,
// plus the caller's pixel offset).
__global__ void handoff_kernel(const uint32_t* __restrict__ match_ij,
                               const uint32_t* __restrict__ match_len,
                               const uint64_t* __restrict__ match_base,
                               const uint64_t* __restrict__ kp_off_a,
                               const uint64_t* __restrict__ kp_off_b, const float2* __restrict__ kp,
                               double px_offset, double2* __restrict__ x1, double2* __restrict__ x2) {
    const uint32_t p = blockIdx.x;
    const uint64_t base = match_base[p];
    const uint32_t len = match_len[p];
    for (uint32_t k = threadIdx.x; k < len; k += blockDim.x) {
        const uint32_t i = match_ij[2 * (base + k)], j = match_ij[2 * (base + k) + 1];
        const float2 a = kp[kp_off_a[p] + i], b = kp[kp_off_b[p] + j];
        x1[base + k] = make_double2((double)a.x + px_offset, (double)a.y + px_offset);
        x2[base + k] = make_double2((double)b.x + px_offset, (double)b.y + px_offset);
    }
}

// Pseudo-random unit rows for the tile tuner: row r, component k from a hash
// of (r, k), then scaled to unit length (one 128-thread block per row).
__global__ void tune_fill_kernel(float* __restrict__ d, uint64_t rows) {
    __shared__ float part[2];
    const uint64_t r = blockIdx.x;
    const uint32_t k = threadIdx.x;
    uint64_t h = (r * 128 + k + 1) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 27;
    const float v = (float)((h >> 40) & 0xffffff) * 0x1p-24f - 0.5f;
    float sq = v * v;
    for (int off = 32; off >= 1; off >>= 1) sq += __shfl_xor(sq, off);
    if ((k & 63) == 0) part[k >> 6] = sq;
    __syncthreads();
    if (r < rows) d[r * 128 + k] = v / sqrtf(part[0] + part[1]);
}

}  // namespace

struct Plan {
    Ctx* ctx = nullptr;
    skh_batch_desc desc{};
    std::vector<uint64_t> view_off;
    std::vector<uint32_t> pairs;
    std::vector<uint64_t> match_base;  // n_pairs + 1
    int tile = kTile256x64;  // the fixed tile, or the tuned plan's most used one
    int forced_tile = -1;    // set by the tuner: every pair on this tile
    uint32_t tbits = 1;
    uint32_t tbits_g = 1;  // FP8 plans: tile bits of the BF16 re-rank (64-column chunks)
    // Row blocks of the ranking launch, grouped by tile: group g holds the
    // blocks [grp_off[g], grp_off[g] + grp_n[g]) of d_blocks (the exact plan
    // uses group 0). The groups launch one after another on the plan's
    // stream: forked onto side streams they would share the process's few
    // hardware queues with other plans' long RANSAC launches and wait behind
    // them, which costs more than the overlap of small groups gains.
    uint32_t grp_off[kNumTiles] = {}, grp_n[kNumTiles] = {}, pairs_per_tile[kNumTiles] = {};
    int grp_tile[kNumTiles] = {};
    int n_groups = 0;
    bool lowp = false;
    bool fp8 = false;
    uint32_t n_blocks = 0;
    uint64_t total_rows = 0, total_cols = 0, total_parts = 0, cap_total = 0;
    uint32_t max_view_rows = 0;

    DevBuf d_view_off, d_tasks_exact, d_tasks_lowp, d_blocks, d_stats, d_q, d_nrm, d_q8, d_meta;
    DevBuf d_row_k1, d_row_j1, d_row_k2, d_row_d1, d_row_d2, d_col_key, d_col_part, d_col_top2;
    DevBuf d_row_match, d_queue, d_queue_len, d_pre, d_chunks, d_match_ij, d_match_len, d_match_base, d_kp_a,
        d_kp_b;
    DevBuf d_x1, d_x2, d_inl, d_F, d_found, d_sel, d_scratch, d_rtasks, d_iters, d_prof, d_order,
        d_est, d_rqueue, d_ver_ij, d_ver_len;
    // FP8 plans: the BF16 form of the views and tasks for the re-rank of open rows
    DevBuf d_stats_b, d_tasks_b;

    hipGraph_t g_match = nullptr, g_ransac = nullptr;
    hipGraphExec_t x_match = nullptr, x_ransac = nullptr;
    // ev[0..2]: run start, matching end, RANSAC end (recorded between graph
    // launches); ev[3..4]: around the distance kernel, recorded inside the graph.
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    bool timing = false;
    bool timed_run = false;

    std::vector<DevBuf*> bufs() {
        return {&d_view_off, &d_tasks_exact, &d_tasks_lowp, &d_blocks, &d_stats,  &d_q,
                &d_nrm,      &d_row_k1,      &d_row_j1,     &d_row_k2, &d_row_d1, &d_row_d2,
                &d_col_key,  &d_col_part,    &d_col_top2,   &d_row_match, &d_queue, &d_queue_len, &d_pre, &d_chunks,
                &d_match_ij, &d_match_len,   &d_match_base, &d_kp_a,   &d_kp_b,   &d_x1,
                &d_x2,       &d_inl,         &d_F,          &d_found,  &d_sel,    &d_scratch,
                &d_rtasks,   &d_iters,       &d_prof,      &d_order,  &d_est,    &d_rqueue,
                &d_q8,       &d_meta,        &d_ver_ij,     &d_ver_len, &d_stats_b, &d_tasks_b};
    }

    ~Plan() {
        if (x_match) (void)hipGraphExecDestroy(x_match);
        if (x_ransac) (void)hipGraphExecDestroy(x_ransac);
        if (g_match) (void)hipGraphDestroy(g_match);
        if (g_ransac) (void)hipGraphDestroy(g_ransac);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        for (DevBuf* b : bufs())
            if (b->ptr) (void)hipFree(b->ptr);
    }

    int alloc(DevBuf& b, size_t bytes, const char* what) {
        bytes = bytes ? bytes : 16;
        hipError_t e = hipMalloc(&b.ptr, bytes);
        if (e != hipSuccess) return ctx->hip(e, what);
        b.cap = bytes;
        return SKH_OK;
    }

    int upload(DevBuf& b, const void* src, size_t bytes, const char* what) {
        int rc = alloc(b, bytes, what);
        if (rc) return rc;
        if (!bytes) return SKH_OK;
        return ctx->hip(hipMemcpy(b.ptr, src, bytes, hipMemcpyHostToDevice), what);
    }

    uint32_t rows(uint32_t v) const { return (uint32_t)(view_off[v + 1] - view_off[v]); }

    int build();
    void launch_distance(hipStream_t s);
    void enqueue_match(hipStream_t s);
    void enqueue_ransac(hipStream_t s);
};

int Plan::build() {
    const uint32_t P = desc.n_pairs, V = desc.n_views;
    fp8 = desc.precision == SKH_PREC_FP8;
    lowp = desc.precision == SKH_PREC_BF16 || fp8;
    for (uint32_t v = 0; v < V; ++v) max_view_rows = std::max(max_view_rows, rows(v));
    if (lowp && max_view_rows > 131072) lowp = fp8 = false;  // tile ids would erode the key precision
    // Tile of every pair: the tuner's forced tile, the context's fixed tile, the
    // context's tuned table for this precision (by size class), or 256 x 64.
    std::vector<int> ptile(P, kTile256x64);
    for (uint32_t p = 0; p < P; ++p) {
        const uint32_t na = rows(pairs[2 * p]), nb = rows(pairs[2 * p + 1]);
        ptile[p] = forced_tile >= 0 ? forced_tile : ctx->pick_tile(desc.precision, na, nb);
    }
    tbits = 1;
    for (uint32_t p = 0; p < P; ++p) {
        pairs_per_tile[ptile[p]]++;
        if (lowp) tbits = std::max(tbits, tile_bits(max_view_rows, tile_bn(ptile[p])));
    }
    tile = (int)(std::max_element(pairs_per_tile, pairs_per_tile + kNumTiles) - pairs_per_tile);
    auto block_rows = [&](uint32_t p) { return lowp ? lowp_block_rows(ptile[p]) : 64u; };

    // per-pair bases
    std::vector<PairTask> pt(P);
    std::vector<LowpTask> lt(P);
    std::vector<uint64_t> kp_a(P), kp_b(P);
    match_base.assign(P + 1, 0);
    uint64_t rb = 0, cb = 0, pb = 0;
    for (uint32_t p = 0; p < P; ++p) {
        const uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
        const uint32_t na = rows(i), nb = rows(j);
        const float* a = desc.d_desc + view_off[i] * kDescLen;
        const float* b = desc.d_desc + view_off[j] * kDescLen;
        pt[p] = PairTask{a, b, na, nb, (uint32_t)rb, (uint32_t)cb};
        LowpTask& l = lt[p];
        l.a = a;
        l.b = b;
        l.na = na;
        l.nb = nb;
        l.row_base = (uint32_t)rb;
        l.col_base = (uint32_t)cb;
        l.part_base = pb;
        l.bm = block_rows(p);
        kp_a[p] = view_off[i];
        kp_b[p] = view_off[j];
        rb += na;
        cb += nb;
        pb += (uint64_t)((na + l.bm - 1) / l.bm) * nb;
        const uint32_t cap = desc.mutual ? std::min(na, nb) : na;
        match_base[p + 1] = match_base[p] + cap;
    }
    total_rows = rb;
    total_cols = cb;
    total_parts = pb;
    cap_total = match_base[P];
    if (rb >= 0xffffffffull || cb >= 0xffffffffull)
        return ctx->fail(SKH_EINVAL, "batch too large for 32-bit row/column bases");

    int rc;
    // quantised views (lowp)
    if ((rc = upload(d_view_off, view_off.data(), 8ull * (V + 1), "view_off"))) return rc;
    const uint64_t total_feat = view_off[V];
    if (fp8) {
        if ((rc = alloc(d_q8, kDescLen * total_feat, "q8")) ||
            (rc = alloc(d_meta, sizeof(Fp8Meta) * total_feat, "meta")))
            return rc;
    }
    if (lowp) {  // FP8 plans keep the BF16 form too, for the re-rank of open rows
        if ((rc = alloc(d_q, 2ull * kDescLen * total_feat, "q")) ||
            (rc = alloc(d_nrm, 4ull * total_feat, "nrm")))
            return rc;
    }
    if (lowp) {
        if ((rc = alloc(d_stats, sizeof(ViewStats) * V, "stats"))) return rc;
        for (uint32_t p = 0; p < P; ++p) {
            const uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
            if (fp8) {
                lt[p].q8a = d_q8.as<uint8_t>() + view_off[i] * kDescLen;
                lt[p].q8b = d_q8.as<uint8_t>() + view_off[j] * kDescLen;
                lt[p].meta_a = d_meta.as<Fp8Meta>() + view_off[i];
                lt[p].meta_b = d_meta.as<Fp8Meta>() + view_off[j];
                lt[p].fp8 = 1;
            } else {
                lt[p].qa = d_q.as<uint16_t>() + view_off[i] * kDescLen;
                lt[p].qb = d_q.as<uint16_t>() + view_off[j] * kDescLen;
                lt[p].nrm_a = d_nrm.as<uint32_t>() + view_off[i];
                lt[p].nrm_b = d_nrm.as<uint32_t>() + view_off[j];
            }
            lt[p].sa = d_stats.as<ViewStats>() + i;
            lt[p].sb = d_stats.as<ViewStats>() + j;
        }
    }
    if (fp8) {
        // the BF16 tasks of the re-rank: same pairs and bases, bf16 rows, norms, statistics
        if ((rc = alloc(d_stats_b, sizeof(ViewStats) * V, "stats_b"))) return rc;
        std::vector<LowpTask> ltb(lt);
        for (uint32_t p = 0; p < P; ++p) {
            const uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
            LowpTask& l = ltb[p];
            l.q8a = l.q8b = nullptr;
            l.meta_a = l.meta_b = nullptr;
            l.fp8 = 0;
            l.qa = d_q.as<uint16_t>() + view_off[i] * kDescLen;
            l.qb = d_q.as<uint16_t>() + view_off[j] * kDescLen;
            l.nrm_a = d_nrm.as<uint32_t>() + view_off[i];
            l.nrm_b = d_nrm.as<uint32_t>() + view_off[j];
            l.sa = d_stats_b.as<ViewStats>() + i;
            l.sb = d_stats_b.as<ViewStats>() + j;
        }
        if ((rc = upload(d_tasks_b, ltb.data(), sizeof(LowpTask) * P, "tasks_b"))) return rc;
        tbits_g = tile_bits(max_view_rows, 64);
    }

    // XCD-aware, longest-first row-block order, one list per tile group
    std::vector<uint32_t> order(P);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        return (uint64_t)pt[x].na * pt[x].nb > (uint64_t)pt[y].na * pt[y].nb;
    });
    std::vector<RowBlock> blocks;
    int ngroups = 0;
    for (int tl = 0; tl < kNumTiles; ++tl) {
        if (lowp ? !pairs_per_tile[tl] : tl > 0) continue;
        std::vector<std::vector<RowBlock>> lists(kXcds);
        std::vector<uint64_t> load(kXcds, 0);
        for (uint32_t p : order) {
            if (pt[p].na == 0 || (lowp && ptile[p] != tl)) continue;
            int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[x] += (uint64_t)pt[p].na * std::max<uint32_t>(pt[p].nb, 1);
            for (uint32_t r0 = 0; r0 < pt[p].na; r0 += lt[p].bm) lists[x].push_back(RowBlock{p, r0});
        }
        size_t longest = 0;
        for (auto& l : lists) longest = std::max(longest, l.size());
        grp_off[ngroups] = (uint32_t)blocks.size();
        for (size_t r = 0; r < longest; ++r)
            for (int x = 0; x < kXcds; ++x)
                blocks.push_back(r < lists[x].size() ? lists[x][r] : RowBlock{kNone, 0u});
        grp_n[ngroups] = (uint32_t)blocks.size() - grp_off[ngroups];
        grp_tile[ngroups] = tl;
        ++ngroups;
    }
    n_groups = ngroups;
    n_blocks = (uint32_t)blocks.size();

    if ((rc = upload(d_tasks_exact, pt.data(), sizeof(PairTask) * P, "tasks")) ||
        (rc = upload(d_tasks_lowp, lt.data(), sizeof(LowpTask) * P, "ltasks")) ||
        (rc = upload(d_blocks, blocks.data(), sizeof(RowBlock) * blocks.size(), "blocks")) ||
        (rc = alloc(d_row_j1, 4ull * rb, "row_j1")) ||
        (rc = alloc(d_row_match, 4ull * rb, "row_match")) ||
        (rc = upload(d_match_base, match_base.data(), 8ull * (P + 1), "match_base")) ||
        (rc = alloc(d_match_ij, 8ull * cap_total, "match_ij")) ||
        (rc = alloc(d_match_len, 4ull * P, "match_len")) ||
        (rc = upload(d_kp_a, kp_a.data(), 8ull * P, "kp_a")) ||
        (rc = upload(d_kp_b, kp_b.data(), 8ull * P, "kp_b")))
        return rc;
    if (lowp) {
        if ((rc = alloc(d_row_k1, 4ull * rb, "row_k1")) || (rc = alloc(d_row_k2, 4ull * rb, "row_k2")) ||
            (rc = alloc(d_col_part, 8ull * pb, "col_part")) ||
            (rc = alloc(d_col_top2, 8ull * cb, "col_top2")) ||
            (rc = alloc(d_queue, 16ull * rb, "queue")) || (rc = alloc(d_pre, 16ull * rb, "pre")) ||
            (rc = alloc(d_chunks, 8ull * chunk_capacity(rb, P), "chunks")) ||
            (rc = alloc(d_queue_len, 8, "queue_len")))
            return rc;
    } else {
        if ((rc = alloc(d_row_d1, 4ull * rb, "row_d1")) || (rc = alloc(d_row_d2, 4ull * rb, "row_d2")) ||
            (rc = alloc(d_col_key, 8ull * cb, "col_key")))
            return rc;
    }

    if (desc.run_ransac) {
        std::vector<RansacTask> rt(P);
        std::vector<uint64_t> sbase(P + 1, 0);
        for (uint32_t p = 0; p < P; ++p) {
            const uint32_t cap = (uint32_t)(match_base[p + 1] - match_base[p]);
            sbase[p + 1] = sbase[p] + ransac_scratch_words(std::max<uint32_t>(cap, 1));
        }
        if ((rc = alloc(d_x1, 16ull * cap_total, "x1")) || (rc = alloc(d_x2, 16ull * cap_total, "x2")) ||
            (rc = alloc(d_inl, cap_total, "inliers")) || (rc = alloc(d_F, 72ull * P, "F")) ||
            (rc = alloc(d_found, 4ull * P, "found")) ||
            (rc = alloc(d_sel, sizeof(skh_model_sel) * P, "sel")) ||
            (rc = alloc(d_scratch, 4ull * sbase[P], "scratch")) ||
            (rc = alloc(d_iters, 8ull * P, "iters")) ||
            (rc = alloc(d_prof, 8ull * 11 * P, "prof")) ||
            (rc = alloc(d_order, 4ull * P, "order")) || (rc = alloc(d_est, 4ull * P, "est")) ||
            (rc = alloc(d_rqueue, 8, "rqueue")) ||
            (rc = alloc(d_ver_ij, 8ull * cap_total, "verified_ij")) ||
            (rc = alloc(d_ver_len, 4ull * P, "verified_len")))
            return rc;
        for (uint32_t p = 0; p < P; ++p) {
            RansacTask& t = rt[p];
            t.x1 = d_x1.as<double2>() + match_base[p];
            t.x2 = d_x2.as<double2>() + match_base[p];
            t.n = (uint32_t)(match_base[p + 1] - match_base[p]);
            t.F = d_F.as<double>() + 9ull * p;
            t.inliers = d_inl.as<uint8_t>() + match_base[p];
            t.found = d_found.as<int32_t>() + p;
            t.sel = d_sel.as<skh_model_sel>() + p;
            t.scratch = d_scratch.as<uint32_t>() + sbase[p];
            t.iters_out = d_iters.as<uint64_t>() + p;
            t.n_dev = d_match_len.as<uint32_t>() + p;
            t.prof = d_prof.as<uint64_t>() + 11ull * p;
            // tracks.rs:56-69: a pair with a model writes its verified matches
            // (the inlier matches in match order) beside its match list
            t.pair_ij = d_match_ij.as<uint32_t>() + 2ull * match_base[p];
            t.ver_ij = d_ver_ij.as<uint32_t>() + 2ull * match_base[p];
            t.ver_len = d_ver_len.as<uint32_t>() + p;
        }
        if ((rc = upload(d_rtasks, rt.data(), sizeof(RansacTask) * P, "rtasks"))) return rc;
    }

    for (hipEvent_t& e : ev)
        if ((rc = ctx->hip(hipEventCreate(&e), "event"))) return rc;
    // Capture the two stage graphs.
    hipStream_t s = ctx->stream;
    if ((rc = ctx->hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture")))
        return rc;
    enqueue_match(s);
    if ((rc = ctx->hip(hipStreamEndCapture(s, &g_match), "end capture")) ||
        (rc = ctx->hip(hipGraphInstantiate(&x_match, g_match, nullptr, nullptr, 0), "instantiate")))
        return rc;
    if (desc.run_ransac) {
        if ((rc = ctx->hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture")))
            return rc;
        enqueue_ransac(s);
        if ((rc = ctx->hip(hipStreamEndCapture(s, &g_ransac), "end capture")) ||
            (rc = ctx->hip(hipGraphInstantiate(&x_ransac, g_ransac, nullptr, nullptr, 0),
                           "instantiate")))
            return rc;
    }
    return SKH_OK;
}

void Plan::enqueue_match(hipStream_t s) {
    const uint32_t P = desc.n_pairs;
    if (lowp) {
        LowpOut o{};
        o.row_k1 = d_row_k1.as<float>();
        o.row_j1 = d_row_j1.as<uint32_t>();
        o.row_k2 = d_row_k2.as<float>();
        o.col_part = d_col_part.as<float2>();
        o.col_top2 = d_col_top2.as<float2>();
        o.row_match = d_row_match.as<uint32_t>();
        o.queue = d_queue.as<uint4>();
        o.queue_len = d_queue_len.as<uint32_t>();
        o.pre = d_pre.as<uint4>();
        o.chunks = fp8 ? d_chunks.as<uint2>() : nullptr;
        launch_stats_reset(d_stats.as<ViewStats>(), desc.n_views, s);
        launch_fill_u32(o.queue_len, 2, 0u, s);
        if (fp8) {
            launch_quantize_views_fp8(desc.d_desc, d_view_off.as<uint64_t>(), desc.n_views,
                                      max_view_rows, d_q8.as<uint8_t>(), d_meta.as<Fp8Meta>(),
                                      d_stats.as<ViewStats>(), s);
            launch_stats_reset(d_stats_b.as<ViewStats>(), desc.n_views, s);
            launch_quantize_views(desc.d_desc, d_view_off.as<uint64_t>(), desc.n_views,
                                  max_view_rows, d_q.as<uint16_t>(), d_nrm.as<uint32_t>(),
                                  d_stats_b.as<ViewStats>(), s);
        } else
            launch_quantize_views(desc.d_desc, d_view_off.as<uint64_t>(), desc.n_views,
                                  max_view_rows, d_q.as<uint16_t>(), d_nrm.as<uint32_t>(),
                                  d_stats.as<ViewStats>(), s);
        const LowpTask* lt = d_tasks_lowp.as<LowpTask>();
        launch_distance(s);
        launch_colmerge(lt, P, o, s);
        launch_decide(lt, P, o, desc.ratio, desc.mutual, tbits, s);
        if (fp8)
            launch_rerank_bf16(d_tasks_b.as<LowpTask>(), lt, o, desc.ratio, desc.mutual, tbits,
                               tbits_g, (uint32_t)std::max(64, 8 * ctx->num_cu), s);
        launch_rescore(lt, o, desc.ratio, desc.mutual, tbits, (uint32_t)std::max(64, 16 * ctx->num_cu),
                       s);
    } else {
        launch_fill_u64(d_col_key.as<unsigned long long>(), total_cols, kColEmpty, s);
        launch_distance(s);
        launch_exact_decide(d_tasks_exact.as<PairTask>(), P, d_row_j1.as<uint32_t>(),
                            d_row_d1.as<float>(), d_row_d2.as<float>(),
                            d_col_key.as<unsigned long long>(), desc.ratio, desc.mutual,
                            d_row_match.as<uint32_t>(), s);
    }
    // out_base of the compaction = match_base (uint64 -> the kernel takes uint32 bases)
    launch_compact(d_tasks_exact.as<PairTask>(), P, d_row_match.as<uint32_t>(),
                          d_match_base.as<uint64_t>(), d_match_ij.as<uint32_t>(),
                          d_match_len.as<uint32_t>(), s);
}

// The distance kernel alone: MFMA ranking (lowp) or the exact FP32 state
// kernel. Its inputs (quantised views, reset column keys) come from the
// preceding launches of a run; re-running it on them is idempotent.
void Plan::launch_distance(hipStream_t s) {
    if (lowp) {
        LowpOut o{};
        o.row_k1 = d_row_k1.as<float>();
        o.row_j1 = d_row_j1.as<uint32_t>();
        o.row_k2 = d_row_k2.as<float>();
        o.col_part = d_col_part.as<float2>();
        const LowpTask* lt = d_tasks_lowp.as<LowpTask>();
        const RowBlock* bl = d_blocks.as<RowBlock>();
        for (int g = 0; g < n_groups; ++g)
            launch_mfma_rank(grp_tile[g], fp8, lt, bl + grp_off[g], grp_n[g], o, tbits, s);
    } else {
        launch_exact_state(d_tasks_exact.as<PairTask>(), d_blocks.as<RowBlock>(), n_blocks,
                           d_row_j1.as<uint32_t>(), d_row_d1.as<float>(), d_row_d2.as<float>(),
                           d_col_key.as<unsigned long long>(), s);
    }
}

void Plan::enqueue_ransac(hipStream_t s) {
    const uint32_t P = desc.n_pairs;
    if (!P) return;
    hipLaunchKernelGGL(handoff_kernel, dim3(P), dim3(256), 0, s, d_match_ij.as<uint32_t>(),
                       d_match_len.as<uint32_t>(), d_match_base.as<uint64_t>(),
                       d_kp_a.as<uint64_t>(), d_kp_b.as<uint64_t>(),
                       reinterpret_cast<const float2*>(desc.d_kp), desc.px_offset,
                       d_x1.as<double2>(), d_x2.as<double2>());
    launch_fill_u32(d_ver_len.as<uint32_t>(), P, 0u, s);  // pairs without a model keep 0
    launch_ransac(d_rtasks.as<RansacTask>(), P, desc.ransac, s, d_rqueue.as<uint32_t>(),
                  d_est.as<uint32_t>(), d_order.as<uint32_t>());
}

}  // namespace skh

using namespace skh;

struct skh_plan : skh::Plan {};

extern "C" int skh_plan_create(skh_ctx* ctx, const skh_batch_desc* desc, skh_plan** out) {
    if (!ctx || !desc || !out) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    *out = nullptr;
    if (!desc->view_offset || (!desc->pairs && desc->n_pairs) || !desc->d_desc ||
        (desc->run_ransac && !desc->d_kp) || !skh_precision_supported(desc->precision))
        return ctx->fail(SKH_EINVAL, "skh_plan_create: invalid argument");
    for (uint32_t v = 0; v < desc->n_views; ++v)
        if (desc->view_offset[v + 1] < desc->view_offset[v])
            return ctx->fail(SKH_EINVAL, "skh_plan_create: view offsets must not decrease");
    for (uint32_t p = 0; p < 2 * desc->n_pairs; ++p)
        if (desc->pairs[p] >= desc->n_views)
            return ctx->fail(SKH_EINVAL, "skh_plan_create: pair %u names a missing view", p / 2);
    if (hipSetDevice(ctx->device) != hipSuccess) return SKH_ENODEV;
    skh_plan* plan = new skh_plan();
    plan->ctx = ctx;
    plan->desc = *desc;
    plan->view_off.assign(desc->view_offset, desc->view_offset + desc->n_views + 1);
    plan->pairs.assign(desc->pairs, desc->pairs + 2 * desc->n_pairs);
    plan->desc.view_offset = nullptr;
    plan->desc.pairs = nullptr;
    int rc = plan->build();
    if (rc) {
        delete plan;
        return rc;
    }
    *out = plan;
    return SKH_OK;
}

extern "C" int skh_plan_destroy(skh_plan* plan) {
    if (!plan) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(plan->ctx->mu);
    (void)hipStreamSynchronize(plan->ctx->stream);
    delete plan;
    return SKH_OK;
}

extern "C" int skh_plan_run(skh_plan* plan) {
    if (!plan) return SKH_EINVAL;
    Ctx* ctx = plan->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    int rc;
    hipStream_t s = ctx->stream;
    const bool t = plan->timing;
    if (t && (rc = ctx->hip(hipEventRecord(plan->ev[0], s), "event"))) return rc;
    if ((rc = ctx->hip(hipGraphLaunch(plan->x_match, s), "graph launch"))) return rc;
    if (t && (rc = ctx->hip(hipEventRecord(plan->ev[1], s), "event"))) return rc;
    if (plan->x_ransac && (rc = ctx->hip(hipGraphLaunch(plan->x_ransac, s), "graph launch")))
        return rc;
    if (t && (rc = ctx->hip(hipEventRecord(plan->ev[2], s), "event"))) return rc;
    plan->timed_run = t;
    return SKH_OK;
}

extern "C" int skh_plan_sync(skh_plan* plan) {
    if (!plan) return SKH_EINVAL;
    return plan->ctx->hip(hipStreamSynchronize(plan->ctx->stream), "sync");
}

extern "C" int skh_plan_match_base(skh_plan* plan, uint64_t* match_base) {
    if (!plan || !match_base) return SKH_EINVAL;
    std::memcpy(match_base, plan->match_base.data(), 8 * plan->match_base.size());
    return SKH_OK;
}

extern "C" int skh_plan_results(skh_plan* plan, uint32_t* match_len, uint32_t* match_ij,
                                uint8_t* inliers, double* F, int32_t* found, skh_model_sel* sel) {
    if (!plan) return SKH_EINVAL;
    Ctx* ctx = plan->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    int rc;
    if ((rc = ctx->hip(hipStreamSynchronize(ctx->stream), "sync"))) return rc;
    const uint32_t P = plan->desc.n_pairs;
    auto get = [&](void* dst, const DevBuf& b, size_t bytes, const char* what) -> int {
        if (!dst || !bytes) return SKH_OK;
        if (!b.ptr) return ctx->fail(SKH_EINVAL, "%s: not computed by this plan", what);
        return ctx->hip(hipMemcpy(dst, b.ptr, bytes, hipMemcpyDeviceToHost), what);
    };
    if ((rc = get(match_len, plan->d_match_len, 4ull * P, "match_len")) ||
        (rc = get(match_ij, plan->d_match_ij, 8ull * plan->cap_total, "match_ij")) ||
        (rc = get(inliers, plan->d_inl, plan->cap_total, "inliers")) ||
        (rc = get(F, plan->d_F, 72ull * P, "F")) || (rc = get(found, plan->d_found, 4ull * P, "found")) ||
        (rc = get(sel, plan->d_sel, sizeof(skh_model_sel) * P, "sel")))
        return rc;
    // A pair whose replay found its batch inconsistent reports found = -1: the
    // whole result is then an error, whether or not the caller asked for found.
    if (plan->d_found.ptr && P) {
        std::vector<int32_t> f(P);
        if ((rc = get(f.data(), plan->d_found, 4ull * P, "found"))) return rc;
        for (uint32_t p = 0; p < P; ++p)
            if (f[p] < 0)
                return ctx->fail(SKH_EHIP, "ransac: internal replay inconsistency in pair %u", p);
    }
    return SKH_OK;
}

extern "C" int skh_plan_ransac_profile(skh_plan* plan, uint64_t* iters, uint64_t* phase,
                                       double* ticks_per_ms) {
    if (!plan || !plan->d_prof.ptr) return SKH_EINVAL;
    Ctx* ctx = plan->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    const uint32_t P = plan->desc.n_pairs;
    int rc;
    if ((rc = ctx->hip(hipStreamSynchronize(ctx->stream), "sync"))) return rc;
    if (iters && (rc = ctx->hip(hipMemcpy(iters, plan->d_iters.ptr, 8ull * P, hipMemcpyDeviceToHost),
                                "iters")))
        return rc;
    if (phase && (rc = ctx->hip(hipMemcpy(phase, plan->d_prof.ptr, 88ull * P, hipMemcpyDeviceToHost),
                                "prof")))
        return rc;
    if (ticks_per_ms) {
        int khz = 0;
        if ((rc = ctx->hip(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device),
                           "clock rate")))
            return rc;
        *ticks_per_ms = (double)khz;
    }
    return SKH_OK;
}

extern "C" int skh_plan_lowp_state(skh_plan* plan, float* row_k1, uint32_t* row_j1,
                                   float* row_k2, float* col_top2, double* eps,
                                   uint64_t* queued) {
    if (!plan) return SKH_EINVAL;
    Ctx* ctx = plan->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (!plan->lowp) return ctx->fail(SKH_EINVAL, "skh_plan_lowp_state: not a BF16/FP8 plan");
    int rc;
    hipStream_t s = ctx->stream;
    if ((rc = ctx->hip(hipStreamSynchronize(s), "sync"))) return rc;
    const uint32_t P = plan->desc.n_pairs;
    auto get = [&](void* dst, const void* src, size_t bytes, const char* what) -> int {
        if (!dst || !bytes) return SKH_OK;
        return ctx->hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), what);
    };
    if ((rc = get(row_k1, plan->d_row_k1.ptr, 4ull * plan->total_rows, "row_k1")) ||
        (rc = get(row_j1, plan->d_row_j1.ptr, 4ull * plan->total_rows, "row_j1")) ||
        (rc = get(row_k2, plan->d_row_k2.ptr, 4ull * plan->total_rows, "row_k2")) ||
        (rc = get(col_top2, plan->d_col_top2.ptr, 8ull * plan->total_cols, "col_top2")))
        return rc;
    if (eps && P) {
        double* d_eps = nullptr;
        if ((rc = ctx->hip(hipMalloc(&d_eps, 8ull * P), "eps"))) return rc;
        launch_pair_eps(plan->d_tasks_lowp.as<LowpTask>(), P, plan->tbits, d_eps, s);
        rc = ctx->hip(hipGetLastError(), "eps kernel");
        if (!rc) rc = ctx->hip(hipMemcpyAsync(eps, d_eps, 8ull * P, hipMemcpyDeviceToHost, s), "eps");
        if (!rc) rc = ctx->hip(hipStreamSynchronize(s), "sync");
        (void)hipFree(d_eps);
        if (rc) return rc;
    }
    if (queued) {
        uint32_t q = 0;
        if ((rc = get(&q, plan->d_queue_len.ptr, 4, "queue_len"))) return rc;
        *queued = q;
    }
    return SKH_OK;
}

extern "C" int skh_plan_set_timing(skh_plan* plan, int on) {
    if (!plan) return SKH_EINVAL;
    plan->timing = on != 0;
    return SKH_OK;
}

extern "C" double skh_plan_time_distance(skh_plan* plan, int reps) {
    if (!plan || reps < 1) return -1.0;
    Ctx* ctx = plan->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    if (hipEventRecord(plan->ev[3], s) != hipSuccess) return -1.0;
    for (int r = 0; r < reps; ++r) plan->launch_distance(s);
    if (hipEventRecord(plan->ev[4], s) != hipSuccess) return -1.0;
    float ms = 0.f;
    if (hipEventSynchronize(plan->ev[4]) != hipSuccess ||
        hipEventElapsedTime(&ms, plan->ev[3], plan->ev[4]) != hipSuccess)
        return -1.0;
    return ms / reps;
}

extern "C" double skh_plan_stage_ms(skh_plan* plan, int stage) {
    if (!plan || stage < 0 || stage > 2) return -1.0;
    if (!plan->timed_run) return -1.0;
    float ms = 0.f;
    hipEvent_t a = stage == 2 ? plan->ev[1] : plan->ev[0];
    hipEvent_t b = stage == 1 ? plan->ev[1] : plan->ev[2];
    if (hipEventSynchronize(b) != hipSuccess) return -1.0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.0;
    return ms;
}

extern "C" int skh_plan_verified(skh_plan* plan, uint32_t* verified_len, uint32_t* verified_ij) {
    if (!plan) return SKH_EINVAL;
    Ctx* ctx = plan->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (!plan->d_ver_len.ptr) return ctx->fail(SKH_EINVAL, "skh_plan_verified: plan runs no RANSAC");
    int rc;
    if ((rc = ctx->hip(hipStreamSynchronize(ctx->stream), "sync"))) return rc;
    const uint32_t P = plan->desc.n_pairs;
    if (verified_len && P &&
        (rc = ctx->hip(hipMemcpy(verified_len, plan->d_ver_len.ptr, 4ull * P, hipMemcpyDeviceToHost),
                       "verified_len")))
        return rc;
    if (verified_ij && plan->cap_total &&
        (rc = ctx->hip(hipMemcpy(verified_ij, plan->d_ver_ij.ptr, 8ull * plan->cap_total,
                                 hipMemcpyDeviceToHost),
                       "verified_ij")))
        return rc;
    return SKH_OK;
}

extern "C" int skh_plan_tiles(skh_plan* plan, uint32_t* pairs_per_tile) {
    if (!plan || !pairs_per_tile) return SKH_EINVAL;
    for (int t = 0; t < kNumTiles; ++t) pairs_per_tile[t] = plan->lowp ? plan->pairs_per_tile[t] : 0u;
    return SKH_OK;
}

// Representative rows (and columns) of the size buckets (<= 192, <= 384,
// <= 1024, <= 3072, more).
static const uint32_t kBucketRep[kSizeBuckets] = {160, 320, 768, 2048, 6000};

extern "C" int skh_ctx_autotune(skh_ctx* ctx, int precision, int32_t* table, double* ms) {
    if (!ctx || (precision != SKH_PREC_BF16 && precision != SKH_PREC_FP8)) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess) return SKH_ENODEV;
    constexpr uint32_t kViews = 32;  // distinct a views and b views per class
    const uint64_t max_rows = 2ull * kViews * kBucketRep[kSizeBuckets - 1];
    float* d_desc = nullptr;
    int rc = ctx->hip(hipMalloc(&d_desc, 4ull * kDescLen * max_rows), "tuner descriptors");
    if (rc) return rc;
    hipLaunchKernelGGL(tune_fill_kernel, dim3((unsigned)max_rows), dim3(128), 0, ctx->stream,
                       d_desc, max_rows);
    if ((rc = ctx->hip(hipStreamSynchronize(ctx->stream), "tuner fill"))) {
        (void)hipFree(d_desc);
        return rc;
    }
    const int saved_override = ctx->tile_override;
    ctx->tile_override = -1;
    int32_t tab[kTileClasses];
    double times[kTileClasses * kNumTiles];
    for (int c = 0; c < kTileClasses && !rc; ++c) {
        const uint32_t na = kBucketRep[c / kSizeBuckets], nb = kBucketRep[c % kSizeBuckets];
        // enough pairs for a few milliseconds of ranking, at least two per XCD
        const uint64_t want = 1500000000ull / ((uint64_t)na * nb);
        const uint32_t P = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want, 64), 4096);
        std::vector<uint64_t> off(2 * kViews + 1, 0);
        for (uint32_t v = 0; v < 2 * kViews; ++v) off[v + 1] = off[v] + (v < kViews ? na : nb);
        std::vector<uint32_t> pl(2 * P);
        for (uint32_t p = 0; p < P; ++p) {
            pl[2 * p] = p % kViews;
            pl[2 * p + 1] = kViews + (p * 7 + p / kViews) % kViews;
        }
        double best = 1e30;
        tab[c] = kTile256x64;
        for (int tl = 0; tl < kNumTiles && !rc; ++tl) {
            skh_plan* plan = new skh_plan();
            plan->ctx = ctx;
            skh_batch_desc d{};
            d.n_views = 2 * kViews;
            d.d_desc = d_desc;
            d.n_pairs = P;
            d.ratio = 0.8f;
            d.mutual = 1;
            d.precision = precision;
            d.run_ransac = 0;
            plan->desc = d;
            plan->view_off = off;
            plan->pairs = pl;
            plan->forced_tile = tl;
            rc = plan->build();
            float el = 0.f;
            if (!rc) rc = ctx->hip(hipGraphLaunch(plan->x_match, ctx->stream), "tuner run");
            if (!rc) rc = ctx->hip(hipEventRecord(plan->ev[3], ctx->stream), "event");
            for (int r = 0; r < 3 && !rc; ++r) plan->launch_distance(ctx->stream);
            if (!rc) rc = ctx->hip(hipEventRecord(plan->ev[4], ctx->stream), "event");
            if (!rc) rc = ctx->hip(hipEventSynchronize(plan->ev[4]), "sync");
            if (!rc) rc = ctx->hip(hipEventElapsedTime(&el, plan->ev[3], plan->ev[4]), "elapsed");
            delete plan;
            const double t = el / 3.0;
            times[c * kNumTiles + tl] = t;
            if (!rc && t < best) {
                best = t;
                tab[c] = tl;
            }
        }
    }
    ctx->tile_override = saved_override;
    (void)hipFree(d_desc);
    if (rc) return rc;
    const int pi = precision == SKH_PREC_FP8 ? 1 : 0;
    std::memcpy(ctx->tile_table[pi], tab, sizeof tab);
    ctx->tile_tuned[pi] = true;
    if (table) std::memcpy(table, tab, sizeof tab);
    if (ms) std::memcpy(ms, times, sizeof times);
    return SKH_OK;
}

extern "C" int skh_ctx_set_tile_table(skh_ctx* ctx, int precision, const int32_t* table) {
    if (!ctx || (precision != SKH_PREC_BF16 && precision != SKH_PREC_FP8)) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    const int pi = precision == SKH_PREC_FP8 ? 1 : 0;
    if (!table) {
        ctx->tile_tuned[pi] = false;
        return SKH_OK;
    }
    for (int c = 0; c < kTileClasses; ++c)
        if (table[c] < 0 || table[c] >= kNumTiles)
            return ctx->fail(SKH_EINVAL, "skh_ctx_set_tile_table: entry %d names no tile", c);
    std::memcpy(ctx->tile_table[pi], table, sizeof(int32_t) * kTileClasses);
    ctx->tile_tuned[pi] = true;
    return SKH_OK;
}

extern "C" int skh_ctx_tile_table(skh_ctx* ctx, int precision, int32_t* table, int* tuned) {
    if (!ctx || !table || (precision != SKH_PREC_BF16 && precision != SKH_PREC_FP8))
        return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    const int pi = precision == SKH_PREC_FP8 ? 1 : 0;
    for (int c = 0; c < kTileClasses; ++c)
        table[c] = ctx->pick_tile(precision, kBucketRep[c / kSizeBuckets], kBucketRep[c % kSizeBuckets]);
    if (tuned) *tuned = ctx->tile_tuned[pi] ? 1 : 0;
    return SKH_OK;
}
