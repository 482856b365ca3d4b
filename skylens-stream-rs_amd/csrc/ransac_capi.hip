// C-ABI of the single-pair ransac_fundamental.
#include <cstring>
#include <mutex>

#include "skh_ctx.hpp"
#include "skh_ransac.hpp"
#include "skylens_hip.h"

using namespace skh;

extern "C" uint64_t skh_ctx_last_ransac_iters(skh_ctx* ctx) { return ctx ? ctx->last_iters : 0; }

extern "C" void skh_ransac_cfg_default(skh_ransac_cfg* c) {
    if (!c) return;
    c->threshold_px = 1.5;
    c->min_inlier_ratio = 0.2;
    c->max_iters = 2000;
    c->confidence = 0.999;
    c->seed = 1;
    c->min_inliers = 15;
    c->max_chance_prob = 1e-6;
}

namespace {
// Output block of one RANSAC call in device memory.
struct ROut {
    double F[9];
    skh_model_sel sel;
    int32_t found;
    int32_t pad;
    uint64_t iters;
};
}  // namespace

extern "C" int skh_ransac_fundamental(skh_ctx* ctx, const double* x1, const double* x2, size_t n,
                                      const skh_ransac_cfg* cfg, double* F_rowmajor,
                                      uint8_t* inliers, skh_model_sel* sel, int* found) {
    if (!ctx) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if ((!x1 && n) || (!x2 && n) || !cfg || !F_rowmajor || !found || (!inliers && n) ||
        n >= 0x7fffffffu || cfg->max_iters > 1000000000ull)
        return ctx->fail(SKH_EINVAL, "skh_ransac_fundamental: invalid argument");
    if (hipSetDevice(ctx->device) != hipSuccess) return SKH_ENODEV;
    *found = 0;
    if (sel) std::memset(sel, 0, sizeof *sel);
    if (n < 8) return SKH_OK;  // matching.rs:947
    int rc;
    const uint32_t nn = (uint32_t)n;
    if ((rc = ctx->ensure(ctx->r_x1, 16 * n, "r_x1")) ||
        (rc = ctx->ensure(ctx->r_x2, 16 * n, "r_x2")) ||
        (rc = ctx->ensure(ctx->r_out, sizeof(ROut) + n, "r_out")) ||
        (rc = ctx->ensure(ctx->r_scratch, 4ull * ransac_scratch_words(nn), "r_scratch")) ||
        (rc = ctx->ensure(ctx->r_task, sizeof(RansacTask), "r_task")) ||
        (rc = ctx->ensure(ctx->r_queue, 8, "r_queue")))
        return rc;
    ROut* o = ctx->r_out.as<ROut>();
    RansacTask t{};
    t.x1 = ctx->r_x1.as<double2>();
    t.x2 = ctx->r_x2.as<double2>();
    t.n = nn;
    t.F = o->F;
    t.sel = &o->sel;
    t.found = &o->found;
    t.iters_out = &o->iters;
    t.inliers = reinterpret_cast<uint8_t*>(o + 1);
    t.scratch = ctx->r_scratch.as<uint32_t>();
    if ((rc = ctx->hip(hipMemcpyAsync(ctx->r_x1.ptr, x1, 16 * n, hipMemcpyHostToDevice, ctx->stream),
                       "copy x1")) ||
        (rc = ctx->hip(hipMemcpyAsync(ctx->r_x2.ptr, x2, 16 * n, hipMemcpyHostToDevice, ctx->stream),
                       "copy x2")) ||
        (rc = ctx->hip(hipMemcpyAsync(ctx->r_task.ptr, &t, sizeof t, hipMemcpyHostToDevice,
                                      ctx->stream),
                       "copy task")))
        return rc;
    launch_ransac(ctx->r_task.as<RansacTask>(), 1, *cfg, ctx->stream, ctx->r_queue.as<uint32_t>());
    if ((rc = ctx->hip(hipGetLastError(), "launch ransac"))) return rc;
    ROut h;
    if ((rc = ctx->hip(hipMemcpyAsync(&h, o, sizeof h, hipMemcpyDeviceToHost, ctx->stream),
                       "copy out")) ||
        (rc = ctx->hip(hipMemcpyAsync(inliers, o + 1, n, hipMemcpyDeviceToHost, ctx->stream),
                       "copy inliers")) ||
        (rc = ctx->hip(hipStreamSynchronize(ctx->stream), "sync")))
        return rc;
    if (h.found < 0) return ctx->fail(SKH_EHIP, "ransac: internal replay inconsistency");
    *found = h.found;
    ctx->last_iters = h.iters;
    if (h.found == 1) {
        std::memcpy(F_rowmajor, h.F, sizeof h.F);
        if (sel) *sel = h.sel;
    }
    return SKH_OK;
}
