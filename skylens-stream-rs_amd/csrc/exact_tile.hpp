// The exact FP32 distance sweep shared by the exact ratio_match kernel and the
// low-precision path's gathered re-score: BM = 64 rows of a, staged transposed
// in LDS, against all nb rows of b in BN = 64-column chunks; each of the 256
// threads owns a 4 x 4 micro-tile of distances computed in the reference's
// order (matching.rs:144-155: lane k takes components 8i + k, t = p - r,
// acc += t * t separately rounded, the eight lanes added in order) and keeps
// its four rows' running top-2 over ascending j (matching.rs:182-187).
#pragma once

#include "skh_internal.hpp"

namespace skh {

constexpr int kExactBM = 64;
constexpr int kExactBN = 64;
constexpr int kExactNT = 256;

struct Top2 {
    uint32_t j;
    float d1, d2;
};

// Merge two row states over disjoint column sets: d1 = min with the lower
// index on equal values, d2 = second smallest of the multiset {a1,a2,b1,b2}.
__device__ __forceinline__ Top2 merge_top2(Top2 a, Top2 b) {
    Top2 r;
    const bool b_wins = (b.d1 < a.d1) || (b.d1 == a.d1 && b.j < a.j);
    if (b_wins) {
        r.j = b.j;
        r.d1 = b.d1;
        r.d2 = a.d1 < b.d2 ? a.d1 : b.d2;
    } else {
        r.j = a.j;
        r.d1 = a.d1;
        r.d2 = b.d1 < a.d2 ? b.d1 : a.d2;
    }
    return r;
}

// The sweep over b. As holds the block's rows transposed (As[c][r]; rows past
// `rows` are zero and never reported). After each chunk, chunk_done(s, col0,
// cols) sees the thread's 4 x 4 distances (rows ty*4.., columns col0 + tx*4..)
// with Bs free for reuse once every thread has passed a barrier. On return,
// st[r] is the top-2 of row ty*4 + r over this thread's columns.
template <typename ChunkDone>
__device__ __forceinline__ void exact_sweep(const float (*As)[kExactBM], float (*Bs)[kExactBN],
                                            const float* __restrict__ b, uint32_t nb, Top2 st[4],
                                            ChunkDone chunk_done) {
    const int tid = threadIdx.x;
    const int tx = tid & 15;  // column group: cols tx*4 .. tx*4+3 of a chunk
    const int ty = tid >> 4;  // row group: rows ty*4 .. ty*4+3 of the block
#pragma unroll
    for (int r = 0; r < 4; ++r) st[r] = Top2{kNone, INFINITY, INFINITY};

    for (uint32_t col0 = 0; col0 < nb; col0 += kExactBN) {
        const uint32_t cols = min<uint32_t>(kExactBN, nb - col0);
        __syncthreads();
        for (int e = tid; e < kExactBN * (kDescLen / 4); e += kExactNT) {
            int cc = e / (kDescLen / 4);
            int c4 = e % (kDescLen / 4);
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if ((uint32_t)cc < cols)
                v = *reinterpret_cast<const float4*>(b + (size_t)(col0 + cc) * kDescLen + 4 * c4);
            Bs[4 * c4 + 0][cc] = v.x;
            Bs[4 * c4 + 1][cc] = v.y;
            Bs[4 * c4 + 2][cc] = v.z;
            Bs[4 * c4 + 3][cc] = v.w;
        }
        __syncthreads();

        float s[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) s[r][q] = 0.f;

        // matching.rs:147-154: lane k takes components k, 8+k, ..., 120+k.
        for (int k = 0; k < 8; ++k) {
            float acc[4][4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[r][q] = 0.f;
#pragma unroll 4
            for (int i = 0; i < 16; ++i) {
                const int c = 8 * i + k;
                float4 av = *reinterpret_cast<const float4*>(&As[c][ty * 4]);
                float4 bv = *reinterpret_cast<const float4*>(&Bs[c][tx * 4]);
                float a4[4] = {av.x, av.y, av.z, av.w};
                float b4[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        float d = a4[r] - b4[q];
                        acc[r][q] = acc[r][q] + d * d;
                    }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) s[r][q] = s[r][q] + acc[r][q];
        }

        // Row state over this thread's 4 columns, ascending j.
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t cj = tx * 4 + q;
            if (cj < cols) {
                const uint32_t j = col0 + cj;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    float d = s[r][q];
                    if (d < st[r].d1) {
                        st[r].d2 = st[r].d1;
                        st[r].d1 = d;
                        st[r].j = j;
                    } else if (d < st[r].d2) {
                        st[r].d2 = d;
                    }
                }
            }
        }
        chunk_done(s, col0, cols);
    }
}

// Merges the 16 column groups of each of the thread's rows (lanes tx of one ty
// group); the result is valid in the tx == 0 lane.
__device__ __forceinline__ Top2 exact_row_merge(Top2 m) {
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) {
        Top2 o;
        o.j = __shfl_xor(m.j, off, 16);
        o.d1 = __shfl_xor(m.d1, off, 16);
        o.d2 = __shfl_xor(m.d2, off, 16);
        m = merge_top2(m, o);
    }
    return m;
}

}  // namespace skh
