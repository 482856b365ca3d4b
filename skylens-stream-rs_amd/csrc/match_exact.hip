// Exact FP32 ratio_match on gfx950: the reference's distance formula, row
// top-2 and column winner, bit for bit.
//
// Reference: crates/core/src/matching.rs:144-215.
//   sq_dist (:144-155): lane k accumulates components 8i+k, t = p - r,
//   acc += t*t with separate roundings (never an FMA), lanes summed 0..7 in
//   order. Row state (:182-187): d < d1 shifts, else d < d2 sets d2; the
//   first index wins d1. Column state (:188-203): strict '<', lowest i wins.
//
// Layout: one workgroup = 64 rows of a x all of b, in 64-column chunks held
// transposed in LDS ([component][row|col]). Each of the 256 threads owns a
// 4x4 micro-tile and keeps two registers per distance (running sum and the
// current lane accumulator), so the 8-lane order costs no extra registers.
// VALU-bound by design: 3 FP32 ops per component and pair, which is the
// minimum for this rounding sequence.
#pragma clang fp contract(off)

#include "exact_tile.hpp"
#include "skh_internal.hpp"

namespace skh {
namespace {

constexpr int BM = kExactBM;
constexpr int BN = kExactBN;
constexpr int NT = kExactNT;

__global__ __launch_bounds__(NT) void exact_state_kernel(const PairTask* __restrict__ tasks,
                                                         const RowBlock* __restrict__ blocks,
                                                         uint32_t* __restrict__ row_j1,
                                                         float* __restrict__ row_d1,
                                                         float* __restrict__ row_d2,
                                                         unsigned long long* __restrict__ col_key) {
    __shared__ float As[kDescLen][BM];
    __shared__ float Bs[kDescLen][BN];

    const RowBlock blk = blocks[blockIdx.x];
    if (blk.pair == kNone) return;  // padding of the XCD-interleaved work list
    const PairTask t = tasks[blk.pair];
    const int tid = threadIdx.x;
    const int tx = tid & 15;
    const int ty = tid >> 4;
    const uint32_t rows = min<uint32_t>(BM, t.na - blk.row0);

    // Load the 64 a-rows transposed: As[c][r]. Missing rows read as 0 and are
    // never reported.
    for (int e = tid; e < BM * (kDescLen / 4); e += NT) {
        int r = e / (kDescLen / 4);
        int c4 = e % (kDescLen / 4);
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if ((uint32_t)r < rows)
            v = *reinterpret_cast<const float4*>(t.a + (size_t)(blk.row0 + r) * kDescLen + 4 * c4);
        As[4 * c4 + 0][r] = v.x;
        As[4 * c4 + 1][r] = v.y;
        As[4 * c4 + 2][r] = v.z;
        As[4 * c4 + 3][r] = v.w;
    }

    Top2 st[4];
    // Column winner of each chunk: rows ascending, strict '<' (matching.rs:188-203).
    exact_sweep(As, Bs, t.b, t.nb, st, [&](const float (&s)[4][4], uint32_t col0, uint32_t cols) {
        unsigned long long part[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float bd = INFINITY;
            uint32_t bi = kNone;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t rr = ty * 4 + r;
                if (rr < rows && s[r][q] < bd) {
                    bd = s[r][q];
                    bi = blk.row0 + rr;
                }
            }
            part[q] = bi == kNone ? kColEmpty
                                  : ((unsigned long long)__float_as_uint(bd) << 32) | bi;
        }
        __syncthreads();  // every thread is done reading Bs
        unsigned long long* cpart = reinterpret_cast<unsigned long long*>(&Bs[0][0]);
#pragma unroll
        for (int q = 0; q < 4; ++q) cpart[ty * BN + tx * 4 + q] = part[q];
        __syncthreads();
        if (tid < (int)cols) {
            unsigned long long best = kColEmpty;
            for (int g = 0; g < NT / 16; ++g) {
                unsigned long long v = cpart[g * BN + tid];
                best = v < best ? v : best;
            }
            if (best != kColEmpty) atomicMin(&col_key[t.col_base + col0 + tid], best);
        }
    });

#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const Top2 m = exact_row_merge(st[r]);
        const uint32_t rr = ty * 4 + r;
        if (tx == 0 && rr < rows) {
            const uint32_t g = t.row_base + blk.row0 + rr;
            row_j1[g] = m.j;
            row_d1[g] = m.d1;
            row_d2[g] = m.d2;
        }
    }
}

__global__ void fill_u64_kernel(unsigned long long* p, size_t n, unsigned long long v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// matching.rs:204-214.
__global__ void exact_decide_kernel(const PairTask* __restrict__ tasks, const uint32_t* row_j1,
                                    const float* row_d1, const float* row_d2,
                                    const unsigned long long* col_key, float ratio, int mutual,
                                    uint32_t* row_match) {
    const PairTask t = tasks[blockIdx.x];  // pairs on x: gridDim.y is capped at 65535
    const float r2 = ratio * ratio;
    for (uint32_t i = blockIdx.y * blockDim.x + threadIdx.x; i < t.na; i += gridDim.y * blockDim.x) {
        const uint32_t g = t.row_base + i;
        const uint32_t j = row_j1[g];
        const float d1 = row_d1[g], d2 = row_d2[g];
        bool ok = j != kNone && isfinite(d2) && !(d1 >= r2 * d2);
        if (ok && mutual) ok = (uint32_t)(col_key[t.col_base + j] & 0xffffffffu) == i;
        row_match[g] = ok ? j : kNone;
    }
}

// One workgroup per pair: ballot-based stream compaction in ascending i.
__global__ __launch_bounds__(256) void compact_kernel(const PairTask* __restrict__ tasks,
                                                      const uint32_t* __restrict__ row_match,
                                                      const uint64_t* __restrict__ out_base,
                                                      uint32_t* __restrict__ out_ij,
                                                      uint32_t* __restrict__ out_len) {
    __shared__ uint32_t wave_cnt[4];
    const uint32_t p = blockIdx.x;
    const PairTask t = tasks[p];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t* out = out_ij + 2ull * out_base[p];
    uint32_t total = 0;
    for (uint32_t base = 0; base < t.na; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t j = i < t.na ? row_match[t.row_base + i] : kNone;
        const bool f = j != kNone;
        const unsigned long long m = __ballot(f);
        const uint32_t before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wave_cnt[wave] = __popcll(m);
        __syncthreads();
        uint32_t off = total;
        for (int w = 0; w < wave; ++w) off += wave_cnt[w];
        if (f) {
            out[2 * (off + before)] = i;
            out[2 * (off + before) + 1] = j;
        }
        total += wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
        __syncthreads();
    }
    if (threadIdx.x == 0) out_len[p] = total;
}

}  // namespace

void launch_exact_state(const PairTask* d_tasks, const RowBlock* d_blocks, uint32_t n_blocks,
                        uint32_t* row_j1, float* row_d1, float* row_d2,
                        unsigned long long* col_key, hipStream_t s) {
    if (n_blocks == 0) return;
    hipLaunchKernelGGL(exact_state_kernel, dim3(n_blocks), dim3(NT), 0, s, d_tasks, d_blocks,
                       row_j1, row_d1, row_d2, col_key);
}

void launch_fill_u64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_exact_decide(const PairTask* d_tasks, uint32_t n_pairs, const uint32_t* row_j1,
                         const float* row_d1, const float* row_d2,
                         const unsigned long long* col_key, float ratio, int mutual,
                         uint32_t* row_match, hipStream_t s) {
    if (n_pairs == 0) return;
    hipLaunchKernelGGL(exact_decide_kernel, dim3(n_pairs, 8), dim3(256), 0, s, d_tasks, row_j1,
                       row_d1, row_d2, col_key, ratio, mutual, row_match);
}

void launch_compact(const PairTask* d_tasks, uint32_t n_pairs, const uint32_t* row_match,
                    const uint64_t* out_base, uint32_t* out_ij, uint32_t* out_len,
                    hipStream_t s) {
    if (n_pairs == 0) return;
    hipLaunchKernelGGL(compact_kernel, dim3(n_pairs), dim3(256), 0, s, d_tasks, row_match,
                       out_base, out_ij, out_len);
}

}  // namespace skh
