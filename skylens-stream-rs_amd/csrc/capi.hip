// C-ABI of libskylens_hip (include/skylens_hip.h): context, buffers, and the
// host-side orchestration of the matching kernels.
#include "skylens_hip.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "skh_ctx.hpp"
#include "skh_internal.hpp"

using namespace skh;

namespace skh {

int Ctx::fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    char buf[512];
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    last_error = buf;
    return code;
}

int Ctx::hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return SKH_OK;
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? SKH_ENOMEM : SKH_EHIP, "%s: %s", what,
                hipGetErrorString(e));
}

int Ctx::ensure(DevBuf& b, size_t bytes, const char* what) {
    if (b.cap >= bytes && b.ptr) return SKH_OK;
    if (b.ptr) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip(e, "hipStreamSynchronize");
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.cap = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    want = want + want / 4;  // grow-only with headroom
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) return hip(e, what);
    b.cap = want;
    return SKH_OK;
}

Ctx::~Ctx() {
    destroy_detector(detector);
    for (DevBuf* b : all_bufs()) {
        if (b->ptr) (void)hipFree(b->ptr);
    }
    if (stream) (void)hipStreamDestroy(stream);
}

// Builds the row-block work list of a single-pair problem and launches the
// exact state kernels. Device-resident inputs.
int Ctx::exact_state_single(const float* d_a, size_t na, const float* d_b, size_t nb) {
    int rc;
    const uint32_t nblk = (uint32_t)((na + 63) / 64);
    if ((rc = ensure(tasks, sizeof(PairTask), "tasks")) ||
        (rc = ensure(blocks, sizeof(RowBlock) * (nblk ? nblk : 1), "blocks")) ||
        (rc = ensure(row_j1, sizeof(uint32_t) * (na ? na : 1), "row_j1")) ||
        (rc = ensure(row_d1, sizeof(float) * (na ? na : 1), "row_d1")) ||
        (rc = ensure(row_d2, sizeof(float) * (na ? na : 1), "row_d2")) ||
        (rc = ensure(col_key, sizeof(unsigned long long) * (nb ? nb : 1), "col_key")))
        return rc;
    PairTask t{d_a, d_b, (uint32_t)na, (uint32_t)nb, 0u, 0u};
    h_blocks.resize(nblk);
    for (uint32_t k = 0; k < nblk; ++k) h_blocks[k] = RowBlock{0u, k * 64u};
    if ((rc = hip(hipMemcpyAsync(tasks.ptr, &t, sizeof t, hipMemcpyHostToDevice, stream),
                  "copy tasks")))
        return rc;
    if (nblk && (rc = hip(hipMemcpyAsync(blocks.ptr, h_blocks.data(), sizeof(RowBlock) * nblk,
                                         hipMemcpyHostToDevice, stream),
                          "copy blocks")))
        return rc;
    // `t` and h_blocks are host memory that the next call may reuse.
    if ((rc = hip(hipStreamSynchronize(stream), "sync"))) return rc;
    launch_fill_u64(col_key.as<unsigned long long>(), nb, kColEmpty, stream);
    launch_exact_state(tasks.as<PairTask>(), blocks.as<RowBlock>(), nblk, row_j1.as<uint32_t>(),
                       row_d1.as<float>(), row_d2.as<float>(), col_key.as<unsigned long long>(),
                       stream);
    return hip(hipGetLastError(), "launch exact_state");
}

int Ctx::ratio_match_dev(const float* d_a, size_t na, const float* d_b, size_t nb, float ratio,
                         int mutual, int precision, uint32_t* d_out_ij, uint32_t* d_out_len) {
    int rc;
    if (precision == SKH_PREC_FP32) {
        if ((rc = exact_state_single(d_a, na, d_b, nb))) return rc;
        if ((rc = ensure(row_match, sizeof(uint32_t) * (na ? na : 1), "row_match")) ||
            (rc = ensure(out_base, sizeof(uint64_t), "out_base")))
            return rc;
        launch_fill_u64(reinterpret_cast<unsigned long long*>(out_base.ptr), 1, 0ull, stream);
        launch_exact_decide(tasks.as<PairTask>(), 1, row_j1.as<uint32_t>(), row_d1.as<float>(),
                            row_d2.as<float>(), col_key.as<unsigned long long>(), ratio, mutual,
                            row_match.as<uint32_t>(), stream);
        launch_compact(tasks.as<PairTask>(), 1, row_match.as<uint32_t>(), out_base.as<uint64_t>(),
                       d_out_ij, d_out_len, stream);
        return hip(hipGetLastError(), "launch decide/compact");
    }
    return ratio_match_lowp(d_a, na, d_b, nb, ratio, mutual, precision, d_out_ij, d_out_len);
}

}  // namespace skh

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------

extern "C" int skh_precision_supported(int precision) {
    return precision == SKH_PREC_FP32 || (precision == SKH_PREC_BF16 && skh::kHaveBf16) ||
           (precision == SKH_PREC_FP8 && skh::kHaveFp8);
}

extern "C" int skh_ctx_create(int device, skh_ctx** out) {
    if (!out) return SKH_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || device < 0 || device >= n)
        return SKH_ENODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SKH_ENODEV;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SKH_ENODEV;
    if (hipSetDevice(device) != hipSuccess) return SKH_ENODEV;
    skh_ctx* c = new skh_ctx();
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SKH_EHIP;
    }
    // The allocation-free entry point's scratch: its device address on this
    // device, looked up once per device with the device current.
    if (noalloc_scratch_address(device, &c->noalloc_scratch) != SKH_OK) c->noalloc_scratch = nullptr;
    *out = c;
    return SKH_OK;
}

extern "C" int skh_ctx_destroy(skh_ctx* ctx) {
    if (!ctx) return SKH_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return SKH_OK;
}

extern "C" void* skh_ctx_stream(skh_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int skh_ctx_set_tile(skh_ctx* ctx, int tile) {
    if (!ctx || tile < SKH_TILE_AUTO || tile > SKH_TILE_256x128) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->tile_override = tile;
    return SKH_OK;
}

extern "C" const char* skh_ctx_last_error(skh_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

extern "C" int skh_candidate_pairs(const uint64_t* views, size_t n, uint64_t temporal,
                                   uint64_t cross, uint64_t pow2_max, uint64_t* out_ij, size_t cap,
                                   size_t* out_len) {
    if ((!views && n) || !out_len) return SKH_EINVAL;
    size_t cnt = 0;
    // matching.rs:25-38: i ascending, j > i ascending.
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = i + 1; j < n; ++j) {
            const uint64_t ca = views[2 * i], pa = views[2 * i + 1];
            const uint64_t cb = views[2 * j], pb = views[2 * j + 1];
            const uint64_t d = pa > pb ? pa - pb : pb - pa;
            bool ok;
            if (ca == cb) {
                const bool pow2 = d != 0 && (d & (d - 1)) == 0;
                ok = d >= 1 && (d <= temporal || (pow2 && d <= pow2_max));
            } else {
                ok = d <= cross;
            }
            if (ok) {
                if (out_ij && cnt < cap) {
                    out_ij[2 * cnt] = i;
                    out_ij[2 * cnt + 1] = j;
                }
                ++cnt;
            }
        }
    }
    *out_len = cnt;
    return cnt > cap ? SKH_EINVAL : SKH_OK;
}

extern "C" void skh_pair_schedule_default(skh_pair_schedule* sch) {
    if (!sch) return;
    *sch = skh_pair_schedule{};
    sch->temporal = 5;  // matching.rs:7-12 PAIR_TEMPORAL, PAIR_POW2_MAX
    sch->pow2_max = 16;
    sch->cross_kind = SKH_CROSS_FORMATION;  // CrossSchedule::FORMATION, :62-71
    sch->right_min = 20;
    sch->left_min = 20;
    sch->max = 40;
    sch->step = 4;
}

extern "C" void skh_pair_schedule_spec(skh_pair_schedule* sch) {
    if (!sch) return;
    skh_pair_schedule_default(sch);
    sch->cross_kind = SKH_CROSS_SPEC;
    sch->range = 4;  // PAIR_CROSS
}

extern "C" int skh_scheduled_pairs(const uint64_t* views, size_t n, const skh_pair_schedule* sch,
                                   uint64_t* out_ij, size_t cap, size_t* out_len) {
    if ((!views && n) || !sch || !out_len ||
        (sch->cross_kind != SKH_CROSS_SPEC && sch->cross_kind != SKH_CROSS_FORMATION))
        return SKH_EINVAL;
    // The reference tests every ordered (i, j) and keeps (min, max), sorted and
    // deduplicated: an unordered pair {i < j} is in iff either direction
    // passes, and walking i < j in order yields the sorted list directly.
    auto pass = [&](size_t i, size_t j) {
        const uint64_t ca = views[2 * i], pa = views[2 * i + 1];
        const uint64_t cb = views[2 * j], pb = views[2 * j + 1];
        const uint64_t d = pa > pb ? pa - pb : pb - pa;
        if (ca == cb) {
            const bool pow2 = d != 0 && (d & (d - 1)) == 0;
            return j > i && d >= 1 && (d <= sch->temporal || (pow2 && d <= sch->pow2_max));
        }
        if (sch->cross_kind == SKH_CROSS_SPEC) return j > i && d <= sch->range;
        uint64_t lo;
        if (ca == 0 && cb == 1)  // CAM_FRONT -> CAM_RIGHT (matching.rs:41-44)
            lo = sch->right_min;
        else if (ca == 0 && cb == 2)  // CAM_FRONT -> CAM_LEFT
            lo = sch->left_min;
        else
            return false;
        const uint64_t step = sch->step > 1 ? sch->step : 1;
        return pb >= pa + lo && pb <= pa + sch->max && (pb - pa - lo) % step == 0;
    };
    size_t cnt = 0;
    for (size_t i = 0; i < n; ++i)
        for (size_t j = i + 1; j < n; ++j)
            if (pass(i, j) || pass(j, i)) {
                if (out_ij && cnt < cap) {
                    out_ij[2 * cnt] = i;
                    out_ij[2 * cnt + 1] = j;
                }
                ++cnt;
            }
    *out_len = cnt;
    return cnt > cap ? SKH_EINVAL : SKH_OK;
}

extern "C" int skh_verified_matches(const uint32_t* matches, size_t n_matches,
                                    const uint8_t* inliers, size_t n_inliers, uint32_t* out_ij,
                                    size_t* out_len) {
    if (!out_len || (n_matches && (!matches || !inliers || !out_ij))) return SKH_EINVAL;
    if (n_matches != n_inliers) return SKH_EINVAL;  // Err((matches.len(), inliers.len()))
    size_t k = 0;
    for (size_t e = 0; e < n_matches; ++e)
        if (inliers[e]) {
            out_ij[2 * k] = matches[2 * e];
            out_ij[2 * k + 1] = matches[2 * e + 1];
            ++k;
        }
    *out_len = k;
    return SKH_OK;
}

static bool valid_precision(int p) { return skh_precision_supported(p) != 0; }

extern "C" int skh_ratio_match_dev(skh_ctx* ctx, const float* d_a, size_t na, const float* d_b,
                                   size_t nb, float ratio, int mutual, int precision,
                                   uint32_t* d_out_ij, uint32_t* d_out_len) {
    if (!ctx) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if ((!d_a && na) || (!d_b && nb) || !d_out_len || (!d_out_ij && na) ||
        !valid_precision(precision) || na >= 0xffffffffu || nb >= 0xffffffffu)
        return ctx->fail(SKH_EINVAL, "skh_ratio_match_dev: invalid argument");
    if (hipSetDevice(ctx->device) != hipSuccess) return SKH_ENODEV;
    return ctx->ratio_match_dev(d_a, na, d_b, nb, ratio, mutual != 0, precision, d_out_ij,
                                d_out_len);
}

extern "C" int skh_ratio_match(skh_ctx* ctx, const float* a, size_t na, const float* b,
                               size_t nb, float ratio, int mutual, int precision,
                               uint32_t* out_ij, size_t cap, size_t* out_len) {
    if (!ctx) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if ((!a && na) || (!b && nb) || !out_len || (!out_ij && cap) || !valid_precision(precision) ||
        na >= 0xffffffffu || nb >= 0xffffffffu)
        return ctx->fail(SKH_EINVAL, "skh_ratio_match: invalid argument");
    if (hipSetDevice(ctx->device) != hipSuccess) return SKH_ENODEV;
    *out_len = 0;
    int rc;
    const size_t bytes_a = sizeof(float) * kDescLen * na, bytes_b = sizeof(float) * kDescLen * nb;
    if ((rc = ctx->ensure(ctx->in_a, bytes_a ? bytes_a : 4, "in_a")) ||
        (rc = ctx->ensure(ctx->in_b, bytes_b ? bytes_b : 4, "in_b")) ||
        (rc = ctx->ensure(ctx->out_ij, sizeof(uint32_t) * 2 * (na ? na : 1), "out_ij")) ||
        (rc = ctx->ensure(ctx->out_len, sizeof(uint32_t), "out_len")))
        return rc;
    if (na && (rc = ctx->hip(hipMemcpyAsync(ctx->in_a.ptr, a, bytes_a, hipMemcpyHostToDevice,
                                            ctx->stream),
                             "copy a")))
        return rc;
    if (nb && (rc = ctx->hip(hipMemcpyAsync(ctx->in_b.ptr, b, bytes_b, hipMemcpyHostToDevice,
                                            ctx->stream),
                             "copy b")))
        return rc;
    if ((rc = ctx->ratio_match_dev(ctx->in_a.as<float>(), na, ctx->in_b.as<float>(), nb, ratio,
                                   mutual != 0, precision, ctx->out_ij.as<uint32_t>(),
                                   ctx->out_len.as<uint32_t>())))
        return rc;
    uint32_t len = 0;
    if ((rc = ctx->hip(hipMemcpyAsync(&len, ctx->out_len.ptr, sizeof len, hipMemcpyDeviceToHost,
                                      ctx->stream),
                       "copy len")) ||
        (rc = ctx->hip(hipStreamSynchronize(ctx->stream), "sync")))
        return rc;
    *out_len = len;
    if (len > cap) return ctx->fail(SKH_EINVAL, "skh_ratio_match: cap %zu < %u matches", cap, len);
    if (len && (rc = ctx->hip(hipMemcpy(out_ij, ctx->out_ij.ptr, sizeof(uint32_t) * 2 * len,
                                        hipMemcpyDeviceToHost),
                              "copy matches")))
        return rc;
    return SKH_OK;
}

extern "C" int skh_ratio_match_state(skh_ctx* ctx, const float* a, size_t na, const float* b,
                                     size_t nb, uint32_t* row_j1, float* row_d1, float* row_d2,
                                     uint32_t* col_i, float* col_d) {
    if (!ctx) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if ((!a && na) || (!b && nb) || (na && (!row_j1 || !row_d1 || !row_d2)) ||
        (nb && (!col_i || !col_d)) || na >= 0xffffffffu || nb >= 0xffffffffu)
        return ctx->fail(SKH_EINVAL, "skh_ratio_match_state: invalid argument");
    if (hipSetDevice(ctx->device) != hipSuccess) return SKH_ENODEV;
    int rc;
    const size_t bytes_a = sizeof(float) * kDescLen * na, bytes_b = sizeof(float) * kDescLen * nb;
    if ((rc = ctx->ensure(ctx->in_a, bytes_a ? bytes_a : 4, "in_a")) ||
        (rc = ctx->ensure(ctx->in_b, bytes_b ? bytes_b : 4, "in_b")))
        return rc;
    if (na && (rc = ctx->hip(hipMemcpyAsync(ctx->in_a.ptr, a, bytes_a, hipMemcpyHostToDevice,
                                            ctx->stream),
                             "copy a")))
        return rc;
    if (nb && (rc = ctx->hip(hipMemcpyAsync(ctx->in_b.ptr, b, bytes_b, hipMemcpyHostToDevice,
                                            ctx->stream),
                             "copy b")))
        return rc;
    if ((rc = ctx->exact_state_single(ctx->in_a.as<float>(), na, ctx->in_b.as<float>(), nb)))
        return rc;
    std::vector<unsigned long long> keys(nb);
    if (na) {
        if ((rc = ctx->hip(hipMemcpyAsync(row_j1, ctx->row_j1.ptr, 4 * na, hipMemcpyDeviceToHost,
                                          ctx->stream),
                           "copy j1")) ||
            (rc = ctx->hip(hipMemcpyAsync(row_d1, ctx->row_d1.ptr, 4 * na, hipMemcpyDeviceToHost,
                                          ctx->stream),
                           "copy d1")) ||
            (rc = ctx->hip(hipMemcpyAsync(row_d2, ctx->row_d2.ptr, 4 * na, hipMemcpyDeviceToHost,
                                          ctx->stream),
                           "copy d2")))
            return rc;
    }
    if (nb && (rc = ctx->hip(hipMemcpyAsync(keys.data(), ctx->col_key.ptr, 8 * nb,
                                            hipMemcpyDeviceToHost, ctx->stream),
                             "copy col")))
        return rc;
    if ((rc = ctx->hip(hipStreamSynchronize(ctx->stream), "sync"))) return rc;
    for (size_t j = 0; j < nb; ++j) {
        if (keys[j] == kColEmpty) {
            col_i[j] = kNone;
            col_d[j] = INFINITY;
        } else {
            col_i[j] = (uint32_t)(keys[j] & 0xffffffffu);
            uint32_t bits = (uint32_t)(keys[j] >> 32);
            std::memcpy(&col_d[j], &bits, 4);
        }
    }
    return SKH_OK;
}
