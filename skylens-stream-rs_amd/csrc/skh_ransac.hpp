// Shared declarations of the device ransac_fundamental.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "skylens_hip.h"

namespace skh {

using RansacCfgDev = skh_ransac_cfg;
using RansacSel = skh_model_sel;

// One pair's RANSAC problem. x1/x2: n correspondences (double2, device).
// Outputs: F (row-major, 9 doubles), inliers (n bytes), found (1 = Some,
// 0 = None), sel. scratch: ransac_scratch_words(n) words of device memory.
struct RansacTask {
    const double2* x1;
    const double2* x2;
    uint32_t n;
    uint32_t pad;
    double* F;
    uint8_t* inliers;
    int32_t* found;
    RansacSel* sel;
    uint32_t* scratch;
    uint64_t* iters_out;  // hypotheses drawn (diagnostics), may be null
    const uint32_t* n_dev;  // if set, the device-side count (<= n) replaces n
    uint64_t* prof;         // if set, 11 phase clocks (wall_clock64 ticks)
    // If set, a pair with a model also writes its verified matches: the
    // (i, j) words of pair_ij whose inlier flag is set, in order, to ver_ij,
    // and their count to *ver_len (left untouched otherwise).
    const uint32_t* pair_ij;
    uint32_t* ver_ij;
    uint32_t* ver_len;
};

uint32_t ransac_scratch_words(uint32_t n);
// Runs every task. queue: 2 words of device scratch (the work counters).
// With `est` and `order` (device, n_tasks words of scratch each), a probe of
// every task's first hypothesis batch estimates its cost and the tasks are
// taken costliest first, the costliest 1/32 by whole workgroups.
void launch_ransac(const RansacTask* tasks, uint32_t n_tasks, const RansacCfgDev& cfg,
                   hipStream_t s, uint32_t* queue, uint32_t* est = nullptr,
                   uint32_t* order = nullptr);

}  // namespace skh
