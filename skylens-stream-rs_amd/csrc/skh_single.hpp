// Device buffers of one single-pair ratio_match pass, and the passes that run
// on them (shared by the context's grow-only buffers and the allocation-free
// entry point's fixed scratch).
#pragma once

#include "skh_internal.hpp"
#include "skh_lowp.hpp"

namespace skh {

struct SingleBufs {
    // low precision
    uint16_t* q_a;
    uint16_t* q_b;
    uint32_t* nrm_a;
    uint32_t* nrm_b;
    uint8_t* q8_a;
    uint8_t* q8_b;
    Fp8Meta* meta_a;
    Fp8Meta* meta_b;
    ViewStats* stats;  // 2
    LowpTask* ltask;   // 1
    float* row_k1;
    uint32_t* row_j1;
    float* row_k2;
    float2* col_part;
    float2* col_top2;
    uint4* queue;         // na entries
    uint32_t* queue_len;  // 2
    uint4* pre;           // na entries
    uint2* chunks;        // chunk_capacity(na, 1) entries
    // exact FP32
    float* row_d1;
    float* row_d2;
    unsigned long long* col_key;
    // shared
    PairTask* task;    // 1
    RowBlock* blocks;
    uint32_t* row_match;
    uint64_t* out_base;
};

void launch_single_setup(const LowpTask* lt, const PairTask& pt, uint32_t nblk, uint32_t bm,
                         LowpTask* d_lt, PairTask* d_pt, RowBlock* d_blocks, hipStream_t s);
void enqueue_lowp_single(const SingleBufs& B, const float* d_a, uint32_t na, const float* d_b,
                         uint32_t nb, float ratio, int mutual, bool fp8, int tile, int num_cu,
                         uint32_t* d_out_ij, uint32_t* d_out_len, hipStream_t s);
void enqueue_exact_single(const SingleBufs& B, const float* d_a, uint32_t na, const float* d_b,
                          uint32_t nb, float ratio, int mutual, uint32_t* d_out_ij,
                          uint32_t* d_out_len, hipStream_t s);

}  // namespace skh
