// The context behind skh_ctx: device, stream, grow-only device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "skh_internal.hpp"
#include "skh_lowp.hpp"

namespace skh {

struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

struct Ctx {
    int device = 0;
    int num_cu = 0;
    int tile_override = -1;  // SKH_TILE_AUTO
    // Tuned tile per size class (tile_class), for BF16 [0] and FP8 [1] ranking.
    int tile_table[2][kTileClasses] = {};
    bool tile_tuned[2] = {false, false};
    uint64_t last_iters = 0;  // hypotheses drawn by the last single-pair RANSAC
    void* noalloc_scratch = nullptr;  // device address of the allocation-free entry's scratch
    void* detector = nullptr;         // feature detector state (features.hip), lazily created
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::string last_error;

    // single-call inputs/outputs
    DevBuf in_a, in_b, out_ij, out_len;
    // exact state
    DevBuf tasks, blocks, row_j1, row_d1, row_d2, col_key, row_match, out_base;
    // low-precision state
    DevBuf q_a, q_b, nrm_a, nrm_b, stats, ltasks, row_k1, row_k2, col_part, col_top2, queue,
        queue_len, q8_a, q8_b, meta_a, meta_b, pre, chunks;
    // single-pair RANSAC
    DevBuf r_x1, r_x2, r_out, r_scratch, r_task, r_queue;
    // track builder: the displacement filter's pairs, points, grid keys, flags
    DevBuf t_pairs, t_point_pair, t_pts, t_keys, t_flags;

    std::vector<RowBlock> h_blocks;

    std::vector<DevBuf*> all_bufs() {
        return {&in_a,   &in_b,   &out_ij,    &out_len,  &tasks,    &blocks, &row_j1,
                &row_d1, &row_d2, &col_key,   &row_match, &out_base, &q_a,    &q_b,
                &nrm_a,  &nrm_b,  &stats,     &ltasks,   &row_k1,   &row_k2, &col_part,
                &col_top2, &queue, &queue_len, &q8_a, &q8_b, &meta_a, &meta_b, &pre, &chunks, &r_x1, &r_x2, &r_out,
                &r_scratch, &r_task, &r_queue, &t_pairs, &t_point_pair, &t_pts, &t_keys,
                &t_flags};
    }

    ~Ctx();

    int fail(int code, const char* fmt, ...);
    int hip(hipError_t e, const char* what);
    int ensure(DevBuf& b, size_t bytes, const char* what);
    // Displacement-filter flags (tracks.rs:519-670) of every pair's sorted,
    // deduplicated (fa, fb) against its images' keypoints (x, y pairs).
    int displacement_flags(const std::vector<std::vector<std::array<uint64_t, 2>>>& valid,
                           const std::vector<std::pair<const double*, const double*>>& kps,
                           std::vector<std::vector<uint8_t>>& flags);

    // Ranking tile of a pair: the fixed override, the tuned table, or the default
    // (256 x 64; 128 x 64 for at most 128 rows).
    int pick_tile(int precision, size_t na, size_t nb) const;
    int exact_state_single(const float* d_a, size_t na, const float* d_b, size_t nb);
    int ratio_match_dev(const float* d_a, size_t na, const float* d_b, size_t nb, float ratio,
                        int mutual, int precision, uint32_t* d_out_ij, uint32_t* d_out_len);
    int ratio_match_lowp(const float* d_a, size_t na, const float* d_b, size_t nb, float ratio,
                         int mutual, int precision, uint32_t* d_out_ij, uint32_t* d_out_len);
};

int noalloc_scratch_address(int device, void** out);
void destroy_detector(void* detector);
size_t noalloc_max_rows();

}  // namespace skh

// The opaque C handle is the context itself.
struct skh_ctx : skh::Ctx {};
