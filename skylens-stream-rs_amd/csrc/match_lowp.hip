// Low-precision ratio_match: BF16 or FP8 (e4m3) MFMA ranking with a rigorous
// error bound, then the reference's FP32 formula for every decision the bound
// leaves open.
//
// Reference semantics: crates/core/src/matching.rs:157-215 (ratio_match),
// reproduced exactly; this file only changes how candidates are found.
//
// 1. quantize: per descriptor row, q = bf16(a) (round to nearest even), the
//    f32 norm n = |a|^2 split into bf16 hi + lo, and per-view maxima of |a|,
//    |a - q| and |q| (the error bound's inputs). Rows with non-finite or huge
//    values flag the view "bad": pairs touching it are scored exactly.
// 2. mfma: per workgroup 4 waves x (32*RB) rows of a against all of b in
//    chunks of 32*CB columns. v_mfma_f32_32x32x16_bf16 over K = 144 =
//    128 components + [n_hi, n_lo, -1/2, -1/2] on the a side and
//    [-1/2, -1/2, n_hi, n_lo] on the b side, so the accumulator is
//    -(|a|^2 + |b|^2 - 2 a.b)/2 = -D/2 directly. Each element then costs
//    5 VALU ops: pack the column-tile id t into the low mantissa bits,
//    row top-2 (v_min, v_med3) and column top-2 (v_min, v_med3). The
//    distance matrix never reaches HBM: rows keep a running top-2 per lane,
//    columns leave one (c1, c2) partial per workgroup.
//    FP8 form: per row, q = e4m3(a / 2^e) with 2^e the smallest power of two
//    >= max|a| / 448 (the row's scale; e4m3 precision is the same at every
//    binade, so rounding the scale up to a power of two costs nothing and lets
//    the matrix core apply it: v_mfma_scale_f32_32x32x64_f8f6f4 with E8M0
//    scales). Two K = 64 steps take the 128 components; a third folds the
//    norms in, |a|^2 as three e4m3 terms (hi + mid + lo, 2^-11 relative) times
//    their own E8M0 scale against -1/2 terms, so the accumulator is again
//    -D/2 and the epilogue is the BF16 one.
// 3. colmerge: partials -> per-column approximate top-2 values.
// 4. decide: with eps bounding |2*key - D_ref| for every element of the pair,
//    each row is a certain accept, a certain reject, or queued.
// 5. rescore: queued rows get the exact FP32 row scan (and, for mutual,
//    the exact column scan of their j1) in the reference's order.
#pragma clang fp contract(off)

#include "exact_tile.hpp"
#include "skh_internal.hpp"
#include "skh_lowp.hpp"

namespace skh {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// e4m3 (OCP "fn" form: bias 7, largest 448, subnormals in steps of 2^-9) of
// x with |x| <= 448, rounded to nearest even.
__device__ __forceinline__ uint32_t e4m3_rne(float x) {
    const uint32_t sign = (__float_as_uint(x) >> 24) & 0x80u;
    const float a = fabsf(x);
    uint32_t code;
    if (a < 0x1p-6f) {
        code = (uint32_t)__builtin_rintf(a * 512.0f);  // 0..8 (8 = 2^-6, code 0x08)
    } else {
        uint32_t u = __float_as_uint(a);
        u += 0x7ffffu + ((u >> 20) & 1u);  // round the mantissa to 3 bits
        code = (((u >> 23) - 120u) << 3) | ((u >> 20) & 7u);
    }
    return sign | code;
}
__device__ __forceinline__ float e4m3_to_f(uint32_t c) {
    const uint32_t e = (c >> 3) & 15u, m = c & 7u;
    const float v = e ? __uint_as_float(((e + 120u) << 23) | (m << 20)) : (float)m * 0x1p-9f;
    return (c & 0x80u) ? -v : v;
}
// Smallest e with m <= 448 * 2^e, for finite m > 0 (448 = 1.75 * 2^8, so the
// first guess is off by at most one step; only f32-subnormal m, whose bound
// underflows, take a few more). The loops are bounded all the same.
__device__ __forceinline__ int fp8_scale_exp(float m) {
    int e = ilogbf(m) - 8;
    for (int k = 0; k < 64 && ldexpf(448.0f, e) < m; ++k) ++e;
    for (int k = 0; k < 64 && ldexpf(448.0f, e - 1) >= m; ++k) --e;
    return e;
}

__device__ __forceinline__ uint32_t bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf16_to_f(uint32_t h) { return __uint_as_float(h << 16); }

// (k1, k2) <- the two smallest of {k1, k2, -p} (k1 <= k2): v_med3_f32 then
// v_min_f32, both with a source negation, updating the registers in place.
// fminf / the med3 builtin would first canonicalise their inputs (one more
// v_max per operand); keys of pairs that reach a decision are never NaN, so
// the plain instructions are exact here.
__device__ __forceinline__ void top2_neg(float& k1, float& k2, float p) {
    asm("v_med3_f32 %1, %0, %1, -%2\n\tv_min_f32 %0, %0, -%2" : "+v"(k1), "+v"(k2) : "v"(p));
}

// Rounds a positive float up by a relative 2^-20 so that bounds computed in
// f32 stay upper bounds.
__device__ __forceinline__ float up(float x) { return x * (1.0f + 0x1p-20f) + 0x1p-126f; }

__device__ __forceinline__ void atomic_max_pos(float* p, float v) {
    atomicMax(reinterpret_cast<unsigned int*>(p), __float_as_uint(v));
}

// Largest |a|^2 ranked by MFMA: real keys stay below 2 kMaxSq = 2e28, far
// under the padding key and inside the f32 accumulator's range.
constexpr float kMaxSq = 1e28f;

// A row's contribution to its view's bound statistics (all zero for a bad row
// or a thread without a row).
struct RowStat {
    float norm, res, qnorm, sq;
    uint32_t bad;
};

// Maxima and the bad flag of a wave's rows (every row of one view), then one
// lane updates the view's statistics: one atomic per wave and statistic
// instead of one per row (the maxima do not depend on the order).
__device__ __forceinline__ void commit_stats(ViewStats* st, RowStat r) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        r.norm = fmaxf(r.norm, __shfl_xor(r.norm, off));
        r.res = fmaxf(r.res, __shfl_xor(r.res, off));
        r.qnorm = fmaxf(r.qnorm, __shfl_xor(r.qnorm, off));
        r.sq = fmaxf(r.sq, __shfl_xor(r.sq, off));
        r.bad |= (uint32_t)__shfl_xor((int)r.bad, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (r.bad) atomicOr(&st->bad, 1u);
        if (r.norm > 0.f) atomic_max_pos(&st->max_norm, r.norm);
        if (r.res > 0.f) atomic_max_pos(&st->max_res, r.res);
        if (r.qnorm > 0.f) atomic_max_pos(&st->max_qnorm, r.qnorm);
        if (r.sq > 0.f) atomic_max_pos(&st->max_sq, r.sq);
    }
}

__device__ RowStat quantize_row(const float* __restrict__ x, uint32_t i,
                                uint16_t* __restrict__ q, uint32_t* __restrict__ nrm);

// One thread per row: bf16 quantisation, packed norm, bound statistics.
__global__ void quantize_kernel(const float* __restrict__ x, uint32_t n, uint16_t* __restrict__ q,
                                uint32_t* __restrict__ nrm, ViewStats* __restrict__ st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    RowStat rs{0.f, 0.f, 0.f, 0.f, 0u};
    if (i < n) rs = quantize_row(x, i, q, nrm);
    commit_stats(st, rs);
}

// All views of a scene: blockIdx.x = view, rows [off[v], off[v+1]) in blocks of
// 256 along y, one thread per row. The rows are staged through LDS 32
// components (one 128-byte line per row) at a time, read by eight
// consecutive threads per row, and the bf16 rows leave the same way, so the
// global accesses are whole lines; each thread still takes its row's
// components in order (the same sums, bit for bit, as quantize_row).
__global__ __launch_bounds__(256) void quantize_views_kernel(const float* __restrict__ x,
                                                             const uint64_t* __restrict__ off,
                                                             uint16_t* __restrict__ q,
                                                             uint32_t* __restrict__ nrm,
                                                             ViewStats* __restrict__ st) {
    constexpr int CH = 32;  // components per stage
    __shared__ float xs[256][CH + 1];
    __shared__ uint32_t qs[256][CH / 2 + 1];
    const uint32_t v = blockIdx.x;
    const uint64_t r0 = off[v] + (uint64_t)blockIdx.y * 256u;
    const uint64_t r_end = off[v + 1];
    if (r0 >= r_end) return;  // the whole workgroup: no row of this view here
    const uint32_t nrows = (uint32_t)min((uint64_t)256, r_end - r0);
    const uint32_t t = threadIdx.x;
    const bool mine = t < nrows;
    float s = 0.f, res = 0.f, qq = 0.f;
    bool bad = false;
    for (int c0 = 0; c0 < kDescLen; c0 += CH) {
        // stage: element e = row * 8 + piece (8 float4 pieces per row)
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t e = t + 256u * k, row = e >> 3, piece = e & 7u;
            if (row < nrows) {
                const float4 w = *reinterpret_cast<const float4*>(
                    x + (size_t)(r0 + row) * kDescLen + c0 + 4 * piece);
                xs[row][4 * piece] = w.x;
                xs[row][4 * piece + 1] = w.y;
                xs[row][4 * piece + 2] = w.z;
                xs[row][4 * piece + 3] = w.w;
            }
        }
        __syncthreads();
        if (mine) {
#pragma unroll
            for (int c = 0; c < CH; c += 2) {
                uint32_t h[2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const float val = xs[t][c + e];
                    bad |= !(fabsf(val) <= 1e15f);  // NaN, inf, or large enough to overflow
                    h[e] = bf16_rne(val);
                    const float qf = bf16_to_f(h[e]);
                    const float d = val - qf;  // exact: a rounding residual is representable
                    s = s + val * val;
                    res = res + d * d;
                    qq = qq + qf * qf;
                }
                qs[t][c / 2] = h[0] | (h[1] << 16);
            }
        }
        __syncthreads();
        // write back: each row's 64 bytes of this stage as four 16-byte pieces
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t e = t + 256u * k, row = e >> 2, piece = e & 3u;
            if (row < nrows) {
                const uint4 w = make_uint4(qs[row][4 * piece], qs[row][4 * piece + 1],
                                           qs[row][4 * piece + 2], qs[row][4 * piece + 3]);
                *reinterpret_cast<uint4*>(q + (size_t)(r0 + row) * kDescLen + c0 + 8 * piece) = w;
            }
        }
    }
    RowStat rs{0.f, 0.f, 0.f, 0.f, 0u};
    if (mine) {
        // the tail of quantize_row: norm split, bad rows, statistics
        const uint32_t i = (uint32_t)(r0 + t);
        bad |= !(s <= kMaxSq);
        uint32_t hi = bf16_rne(s);
        const float lo = s - bf16_to_f(hi);
        uint32_t lo16 = bf16_rne(lo);
        if (bad) {
            hi = 0;
            lo16 = 0;
        }
        nrm[i] = hi | (lo16 << 16);
        rs = bad ? RowStat{0.f, 0.f, 0.f, 0.f, 1u}
                 : RowStat{up(sqrtf(s) * 1.0001f), up(sqrtf(res) * 1.0001f),
                           up(sqrtf(qq) * 1.0001f), up(s * 1.0001f), 0u};
    }
    commit_stats(st + v, rs);
}

__device__ RowStat quantize_row(const float* __restrict__ x, uint32_t i,
                                uint16_t* __restrict__ q, uint32_t* __restrict__ nrm) {
    const float* r = x + (size_t)i * kDescLen;
    float s = 0.f, res = 0.f, qq = 0.f;
    bool bad = false;
    uint32_t* qo = reinterpret_cast<uint32_t*>(q + (size_t)i * kDescLen);
    for (int c = 0; c < kDescLen; c += 4) {
        float4 v = *reinterpret_cast<const float4*>(r + c);
        float vv[4] = {v.x, v.y, v.z, v.w};
        uint32_t h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            bad |= !(fabsf(vv[e]) <= 1e15f);  // NaN, inf, or large enough to overflow
            h[e] = bf16_rne(vv[e]);
            float qf = bf16_to_f(h[e]);
            float d = vv[e] - qf;  // exact: a rounding residual is representable
            s = s + vv[e] * vv[e];
            res = res + d * d;
            qq = qq + qf * qf;
        }
        qo[c / 2] = h[0] | (h[1] << 16);
        qo[c / 2 + 1] = h[2] | (h[3] << 16);
    }
    // Padding rows and columns carry the norm kPadNorm (~1e30), so their keys
    // (~5e29) must stay above every real key, which is at most
    // (|a| + |b|)^2 / 2 <= 2 max |a|^2. Rows with |a|^2 above kMaxSq are
    // scored exactly instead.
    bad |= !(s <= kMaxSq);
    uint32_t hi = bf16_rne(s);
    float lo = s - bf16_to_f(hi);
    uint32_t lo16 = bf16_rne(lo);
    if (bad) {
        hi = 0;
        lo16 = 0;
    }
    nrm[i] = hi | (lo16 << 16);
    if (bad) return RowStat{0.f, 0.f, 0.f, 0.f, 1u};
    // |a|, |a - q| and |q| from f32 sums of squares: the sums carry a relative
    // error below 130 * 2^-24, covered by the 1.0001 factor.
    return RowStat{up(sqrtf(s) * 1.0001f), up(sqrtf(res) * 1.0001f), up(sqrtf(qq) * 1.0001f),
                   up(s * 1.0001f), 0u};
}

// FP8 form of quantize_row: e4m3 bytes of a / 2^e, the row's E8M0 scale code
// and the three-term fold of |a|^2, plus the same bound statistics in real
// units (q(a) = e4m3 value * 2^e).
__device__ RowStat quantize_row_fp8(const float* __restrict__ x, uint32_t i,
                                    uint8_t* __restrict__ q, Fp8Meta* __restrict__ meta) {
    const float* r = x + (size_t)i * kDescLen;
    float s = 0.f, mx = 0.f;
    bool bad = false;
    for (int c = 0; c < kDescLen; c += 4) {
        const float4 v = *reinterpret_cast<const float4*>(r + c);
        const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            bad |= !(fabsf(vv[e]) <= 1e15f);
            s = s + vv[e] * vv[e];  // the same f32 sum as the BF16 quantiser
            mx = fmaxf(mx, fabsf(vv[e]));
        }
    }
    bad |= !(s <= kMaxSq);
    // Scale exponents only of finite, bounded rows (|v| <= 1e15, s <= 1e28):
    // there fp8_scale_exp's first guess is within a step or two of the answer.
    int e = 0, f = 0;
    if (!bad) {
        e = mx > 0.f ? fp8_scale_exp(mx) : 0;
        f = s > 0.f ? fp8_scale_exp(s) : 0;
        bad = e < -127 || e > 127 || f < -127 || f > 127;  // outside E8M0
    }
    uint32_t* qo = reinterpret_cast<uint32_t*>(q + (size_t)i * kDescLen);
    if (bad) {
        for (int c = 0; c < kDescLen / 4; ++c) qo[c] = 0u;
        meta[i] = Fp8Meta{127u | (127u << 8), 0u};
        return RowStat{0.f, 0.f, 0.f, 0.f, 1u};
    }
    float res = 0.f, qq = 0.f;
    for (int c = 0; c < kDescLen; c += 4) {
        const float4 v = *reinterpret_cast<const float4*>(r + c);
        const float vv[4] = {v.x, v.y, v.z, v.w};
        uint32_t word = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t b = e4m3_rne(ldexpf(vv[k], -e));
            const float qf = ldexpf(e4m3_to_f(b), e);
            const float d = vv[k] - qf;  // exact: a rounding residual is representable
            res = res + d * d;
            qq = qq + qf * qf;
            word |= b << (8 * k);
        }
        qo[c / 4] = word;
    }
    uint32_t fold = 0u;
    if (s > 0.f) {
        const float xs = ldexpf(s, -f);  // in (224, 448]
        const uint32_t hi = e4m3_rne(xs);
        const float r1 = xs - e4m3_to_f(hi);  // exact (Sterbenz)
        const uint32_t mid = e4m3_rne(r1);
        const float r2 = r1 - e4m3_to_f(mid);
        const uint32_t lo = e4m3_rne(r2);
        fold = hi | (mid << 8) | (lo << 16);
    }
    meta[i] = Fp8Meta{(uint32_t)(e + 127) | ((uint32_t)(f + 127) << 8), fold};
    return RowStat{up(sqrtf(s) * 1.0001f), up(sqrtf(res) * 1.0001f), up(sqrtf(qq) * 1.0001f),
                   up(s * 1.0001f), 0u};
}

__global__ void quantize_fp8_kernel(const float* __restrict__ x, uint32_t n,
                                    uint8_t* __restrict__ q, Fp8Meta* __restrict__ meta,
                                    ViewStats* __restrict__ st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    RowStat rs{0.f, 0.f, 0.f, 0.f, 0u};
    if (i < n) rs = quantize_row_fp8(x, i, q, meta);
    commit_stats(st, rs);
}

__global__ void quantize_views_fp8_kernel(const float* __restrict__ x,
                                          const uint64_t* __restrict__ off,
                                          uint8_t* __restrict__ q, Fp8Meta* __restrict__ meta,
                                          ViewStats* __restrict__ st) {
    const uint32_t v = blockIdx.x;
    const uint64_t r = off[v] + (uint64_t)blockIdx.y * blockDim.x + threadIdx.x;
    RowStat rs{0.f, 0.f, 0.f, 0.f, 0u};
    if (r < off[v + 1]) rs = quantize_row_fp8(x, (uint32_t)r, q, meta);
    commit_stats(st + v, rs);
}

__global__ void stats_reset_kernel(ViewStats* st, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st[i] = ViewStats{0.f, 0.f, 0.f, 0.f, 0u, 0u};
}

// ---------------------------------------------------------------------------
// MFMA ranking kernel
// ---------------------------------------------------------------------------

// a-side fragment of the last K step: lanes 0..31 carry [n_hi, n_lo, -1/2, -1/2].
__device__ __forceinline__ bf16x8 tail_frag(uint32_t nrm_packed, bool a_side, bool lo_half) {
    union {
        uint32_t u[4];
        bf16x8 v;
    } f;
    const uint32_t half2 = 0xbf00bf00u;  // two bf16 -0.5
    f.u[0] = lo_half ? (a_side ? nrm_packed : half2) : 0u;
    f.u[1] = lo_half ? (a_side ? half2 : nrm_packed) : 0u;
    f.u[2] = 0u;
    f.u[3] = 0u;
    return f.v;
}

__device__ __forceinline__ bf16x8 load_frag(const uint16_t* p) {
    union {
        uint4 u;
        bf16x8 v;
    } f;
    f.u = *reinterpret_cast<const uint4*>(p);
    return f.v;
}

// Huge norm used for padding rows/columns: their keys never win a top-2.
constexpr uint32_t kPadNorm = 0x00007149u;  // bf16 hi = 0x7149 (~1e30), lo = 0
// FP8 padding: zero components, fold 448 * 2^92 (~2.2e30).
constexpr uint32_t kPadCodes = 127u | ((127u + 92u) << 8);
constexpr uint32_t kPadFold = 0x7eu;
constexpr uint32_t kNegHalf3 = 0x00b0b0b0u;  // three e4m3 -1/2 bytes

// Loads through the global address space: global_load instructions, which
// only the vector-memory counter tracks. Through a generic pointer they would
// be flat loads, which the LDS counter tracks as well, so the LDS waits of the
// tile loop would also wait for the next chunk's prefetch.
template <class T>
__device__ __forceinline__ T ldg_as(const void* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}
__device__ __forceinline__ Fp8Meta ldg_meta8(const Fp8Meta* p) {
    typedef uint32_t v2u __attribute__((ext_vector_type(2)));
    const v2u v = *(const __attribute__((address_space(1))) v2u*)p;
    return Fp8Meta{v.x, v.y};
}
// 16 bytes as a native vector (a uint4 copy would go through its generic
// copy constructor, a flat load again)
__device__ __forceinline__ uint4 ldg_u4(const void* p) {
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    const v4u v = *(const __attribute__((address_space(1))) v4u*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Operand handling of the two ranking precisions. Per 32 x 32 tile the matrix
// core runs kSteps MFMAs; a row of b takes kRowBytes in global memory and
// kStride bytes in LDS (16 bytes of padding keep the 16-byte reads of 8
// consecutive lanes on distinct banks).
template <bool F8>
struct RankOps;

template <>
struct RankOps<false> {
    static constexpr int kSteps = 9;  // K = 144: 128 components + the norm fold
    static constexpr int kRowBytes = 2 * kDescLen;
    static constexpr int kStride = 272;
    using Meta = uint32_t;  // packed bf16 norm
    struct AFrag {
        bf16x8 f[9];
    };
    struct BFrag {
        bf16x8 f[9];
    };
    static __device__ __forceinline__ Meta pad_meta() { return kPadNorm; }
    static __device__ __forceinline__ Meta meta_a(const LowpTask& t, uint32_t row) {
        return ldg_as<Meta>(t.nrm_a + row);
    }
    static __device__ __forceinline__ Meta meta_b(const LowpTask& t, uint32_t col) {
        return ldg_as<Meta>(t.nrm_b + col);
    }
    static __device__ __forceinline__ const uint8_t* rows_a(const LowpTask& t) {
        return reinterpret_cast<const uint8_t*>(t.qa);
    }
    static __device__ __forceinline__ const uint8_t* rows_b(const LowpTask& t) {
        return reinterpret_cast<const uint8_t*>(t.qb);
    }
    // lane (row l32, half h) holds components 16 s + 8 h .. + 7 of step s
    static __device__ __forceinline__ void load_a(AFrag& a, const uint8_t* row, bool valid,
                                                  Meta m, bool lo_half) {
        const uint16_t* src = reinterpret_cast<const uint16_t*>(row) + 8 * (lo_half ? 0 : 1);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            a.f[s] = load_frag(src + 16 * s);
            if (!valid) a.f[s] = bf16x8{};
        }
        a.f[8] = tail_frag(m, true, lo_half);
    }
    static __device__ __forceinline__ void load_b(BFrag& b, const uint8_t* lds_row, Meta m,
                                                  bool lo_half) {
        const uint16_t* src = reinterpret_cast<const uint16_t*>(lds_row) + 8 * (lo_half ? 0 : 1);
#pragma unroll
        for (int s = 0; s < 8; ++s) b.f[s] = load_frag(src + 16 * s);
        b.f[8] = tail_frag(m, false, lo_half);
    }
    static __device__ __forceinline__ f32x16 mfma(const AFrag& a, const BFrag& b, int s,
                                                  f32x16 acc) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.f[s], b.f[s], acc, 0, 0, 0);
    }
};

// e4m3 operands of v_mfma_scale_f32_32x32x64_f8f6f4. Lane (row l32, half h)
// holds 32 bytes per step; bytes 0-15 of every lane are K block 0 and bytes
// 16-31 block 1, and block b's E8M0 scale of row/column l32 is the scale
// operand of lane l32 + 32 b. Steps 0 and 1 take components 64 s + 32 h ..
// + 31 (A and B alike, so any such assignment gives the dot product), with the
// row's (column's) scale in both halves. Step 2 is the norm fold: the a row's
// [hi, mid, lo] at bytes 0-2 (block 0, the fold scale) against the b column's
// three -1/2 (block 0, scale 1), and three -1/2 at bytes 16-18 (block 1, scale
// 1) against the b column's [hi, mid, lo] (block 1, its fold scale):
// -(|a|^2 + |b|^2)/2 to 2^-11 relative.
template <>
struct RankOps<true> {
    static constexpr int kSteps = 3;
    static constexpr int kRowBytes = kDescLen;
    static constexpr int kStride = 144;
    using Meta = Fp8Meta;
    struct AFrag {
        i32x8 f[3];
        int scale, fold_scale;
    };
    struct BFrag {
        i32x8 f[3];
        int scale, fold_scale;
    };
    static __device__ __forceinline__ Meta pad_meta() { return Fp8Meta{kPadCodes, kPadFold}; }
    static __device__ __forceinline__ Meta meta_a(const LowpTask& t, uint32_t row) {
        return ldg_meta8(t.meta_a + row);
    }
    static __device__ __forceinline__ Meta meta_b(const LowpTask& t, uint32_t col) {
        return ldg_meta8(t.meta_b + col);
    }
    static __device__ __forceinline__ const uint8_t* rows_a(const LowpTask& t) { return t.q8a; }
    static __device__ __forceinline__ const uint8_t* rows_b(const LowpTask& t) { return t.q8b; }
    static __device__ __forceinline__ i32x8 load32(const uint8_t* p) {
        union {
            uint4 u[2];
            i32x8 v;
        } f;
        f.u[0] = *reinterpret_cast<const uint4*>(p);
        f.u[1] = *reinterpret_cast<const uint4*>(p + 16);
        return f.v;
    }
    static __device__ __forceinline__ i32x8 fold_frag(uint32_t w0, uint32_t w4, bool lo_half) {
        i32x8 v = i32x8{};
        if (lo_half) {
            v[0] = (int)w0;
            v[4] = (int)w4;
        }
        return v;
    }
    static __device__ __forceinline__ void load_a(AFrag& a, const uint8_t* row, bool valid, Meta m,
                                                  bool lo_half) {
        const uint8_t* src = row + 32 * (lo_half ? 0 : 1);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            a.f[s] = load32(src + 64 * s);
            if (!valid) a.f[s] = i32x8{};
        }
        a.f[2] = fold_frag(m.fold, kNegHalf3, lo_half);
        a.scale = (int)(m.codes & 0xffu);
        a.fold_scale = lo_half ? (int)((m.codes >> 8) & 0xffu) : 127;
    }
    static __device__ __forceinline__ void load_b(BFrag& b, const uint8_t* lds_row, Meta m,
                                                  bool lo_half) {
        const uint8_t* src = lds_row + 32 * (lo_half ? 0 : 1);
#pragma unroll
        for (int s = 0; s < 2; ++s) b.f[s] = load32(src + 64 * s);
        b.f[2] = fold_frag(kNegHalf3, m.fold, lo_half);
        b.scale = (int)(m.codes & 0xffu);
        b.fold_scale = lo_half ? 127 : (int)((m.codes >> 8) & 0xffu);
    }
    static __device__ __forceinline__ f32x16 mfma(const AFrag& a, const BFrag& b, int s,
                                                  f32x16 acc) {
        if (s < 2)
            return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a.f[s], b.f[s], acc, 0, 0, 0,
                                                                   a.scale, 0, b.scale);
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a.f[2], b.f[2], acc, 0, 0, 0,
                                                               a.fold_scale, 0, b.fold_scale);
    }
};

// The 256 x 128 shape is held to 256 VGPRs (two waves per SIMD, a few spills
// outside the MFMA chains): 12.9 -> 11.7 ms on a 3663-pair scene.
template <bool F8, int RB, int CB>
__global__ __launch_bounds__(256, RB == 2 && CB == 4 ? 2 : 1) void mfma_rank_kernel(
    const LowpTask* __restrict__ tasks, const RowBlock* __restrict__ blocks, LowpOut out,
    uint32_t tbits) {
    using Ops = RankOps<F8>;
    using Meta = typename Ops::Meta;
    constexpr int WR = 32 * RB;         // rows per wave
    constexpr int BM = 4 * WR;          // rows per workgroup
    constexpr int BN = 32 * CB;         // columns per chunk
    constexpr int STRIDE = Ops::kStride;
    constexpr int PIECES = Ops::kRowBytes / 16;  // 16-byte pieces of one row of b
    constexpr int PF = BN * PIECES / 256;        // pieces of a chunk per thread
    constexpr int KS = Ops::kSteps;
    // Two b chunks in LDS: chunk ch + 1 is fetched from global memory into
    // registers while chunk ch is multiplied, then stored into the other
    // buffer; one barrier per chunk. The column partials alternate the same way.
    __shared__ __attribute__((aligned(16))) uint8_t Bs[2][BN * STRIDE];
    __shared__ Meta Bn[2][BN];
    __shared__ float2 Cs[2][4][BN];

    const RowBlock blk = blocks[blockIdx.x];
    if (blk.pair == kNone) return;  // padding of the XCD-interleaved work list
    const LowpTask t = tasks[blk.pair];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int l32 = lane & 31;
    const bool lo_half = lane < 32;
    const uint32_t tmask = (1u << tbits) - 1u;
    uint32_t keep_v;  // ~tmask, pinned to a VGPR
    asm volatile("v_mov_b32 %0, %1" : "=v"(keep_v) : "s"(~tmask));

    // a fragments held in registers for the whole sweep.
    typename Ops::AFrag afrag[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const uint32_t row = blk.row0 + wave * WR + rb * 32 + l32;
        const bool valid = row < t.na;
        const uint8_t* src = Ops::rows_a(t) + (size_t)(valid ? row : 0) * Ops::kRowBytes;
        Ops::load_a(afrag[rb], src, valid, valid ? Ops::meta_a(t, row) : Ops::pad_meta(), lo_half);
    }

    // Running row top-2 per accumulator register: m1 = packed key, m2 = value.
    float m1[RB][16], m2[RB][16];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            m1[rb][r] = INFINITY;
            m2[rb][r] = INFINITY;
        }

    // chunk loader: thread tid fetches pieces e = tid + 256 k (column
    // e / PIECES, 16-byte segment e % PIECES) and, for tid < BN, the column's
    // metadata
    uint4 pf[PF];
    Meta pfn = Ops::pad_meta();
    const uint8_t* qb = Ops::rows_b(t);
    auto fetch = [&](uint32_t col0) {
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int e = tid + 256 * k, c = e / PIECES, seg = e % PIECES;
            const uint32_t col = col0 + c;
            pf[k] = col < t.nb ? ldg_u4(qb + (size_t)col * Ops::kRowBytes + 16 * seg)
                               : make_uint4(0, 0, 0, 0);
        }
        if (tid < BN) pfn = (col0 + tid < t.nb) ? Ops::meta_b(t, col0 + tid) : Ops::pad_meta();
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int e = tid + 256 * k, c = e / PIECES, seg = e % PIECES;
            *reinterpret_cast<uint4*>(&Bs[buf][c * STRIDE + 16 * seg]) = pf[k];
        }
        if (tid < BN) Bn[buf][tid] = pfn;
    };
    // merge of the 4 waves' column partials of chunk `ch` (buffer ch & 1)
    auto merge_cols = [&](uint32_t ch) {
        if (tid < BN) {
            const uint32_t col = ch * BN + tid;
            const int cb = ch & 1;
            if (col < t.nb) {
                float a1 = Cs[cb][0][tid].x, a2 = Cs[cb][0][tid].y;
#pragma unroll
                for (int w = 1; w < 4; ++w) {
                    float b1 = Cs[cb][w][tid].x, b2 = Cs[cb][w][tid].y;
                    float n2 = fminf(fmaxf(a1, b1), fminf(a2, b2));
                    a1 = fminf(a1, b1);
                    a2 = n2;
                }
                out.col_part[(size_t)t.part_base + (size_t)(blk.row0 / BM) * t.nb + col] =
                    make_float2(a1, a2);
            }
        }
    };

    const uint32_t nchunks = (t.nb + BN - 1) / BN;
    if (nchunks > 0) {
        fetch(0);
        store(0);
    }
    __syncthreads();
    // b fragments of column tile cb of LDS buffer buf
    auto load_b = [&](typename Ops::BFrag& bf, int buf, int cb) {
        Ops::load_b(bf, &Bs[buf][(cb * 32 + l32) * STRIDE], Bn[buf][cb * 32 + l32], lo_half);
    };
    // epilogue of accumulator registers r0..r1-1 of row tile rb: key = -(acc with
    // the tile id in its low bits) ~ D/2, one v_and_or_b32 (keep_v is a VGPR so
    // the SGPR tile id fits gfx9's constant bus), then row and column top-2.
    auto epilogue = [&](const f32x16& acc, int rb, int r0, int r1, uint32_t tile, float& c1,
                        float& c2) {
#pragma unroll
        for (int r = r0; r < r1; ++r) {
            const float p = __uint_as_float((__float_as_uint(acc[r]) & keep_v) | tile);
            top2_neg(m1[rb][r], m2[rb][r], p);
            top2_neg(c1, c2, p);
        }
    };
    // Column partial of tile cb: merge the two half-waves (same column, other
    // 16*RB rows), one (c1, c2) per wave and column into Cs.
    auto column_out = [&](float c1, float c2, int buf, int cb) {
        const float o1 = __shfl_xor(c1, 32);
        const float o2 = __shfl_xor(c2, 32);
        const float n2 = fminf(fmaxf(c1, o1), fminf(c2, o2));
        c1 = fminf(c1, o1);
        c2 = n2;
        if (lo_half) Cs[buf][wave][cb * 32 + l32] = make_float2(c1, c2);
    };

    // Software pipeline over the CB*RB tiles of a chunk (tile ti = (cb, rb) =
    // (ti / RB, ti % RB)): the MFMAs of tile ti + 1 are issued one at a time,
    // each followed by a slice of tile ti's epilogue, so the VALU work runs
    // while the matrix core is busy. The scheduling barriers keep that
    // interleave; the top-2 updates per register happen in the same order as a
    // plain tile-by-tile sweep, so the results do not depend on it. With two
    // tiles per chunk (128 x 64) the second accumulator costs the kernel its
    // third wave per SIMD, which is worth more: that shape runs the plain sweep.
    constexpr bool kPipe = CB * RB > 2;
    for (uint32_t ch = 0; ch < nchunks; ++ch) {
        const int buf = ch & 1;
        if (ch + 1 < nchunks) fetch((ch + 1) * BN);
        if (ch > 0) merge_cols(ch - 1);

        typename Ops::BFrag bfrag;
        load_b(bfrag, buf, 0);
        f32x16 acc[2];
        if (kPipe) {
            acc[0] = f32x16{};
#pragma unroll
            for (int s = 0; s < KS; ++s) acc[0] = Ops::mfma(afrag[0], bfrag, s, acc[0]);
        }
        float c1 = INFINITY, c2 = INFINITY;
#pragma unroll
        for (int ti = 0; ti < CB * RB; ++ti) {
            const int cb = ti / RB, rb = ti % RB, cur = ti & 1;
            const uint32_t tile = (ch * CB + cb) & tmask;
            if (!kPipe) {
                if (rb == 0 && ti > 0) load_b(bfrag, buf, cb);
                acc[cur] = f32x16{};
#pragma unroll
                for (int s = 0; s < KS; ++s) acc[cur] = Ops::mfma(afrag[rb], bfrag, s, acc[cur]);
                epilogue(acc[cur], rb, 0, 16, tile, c1, c2);
            } else if (ti + 1 < CB * RB) {
                const int nrb = (ti + 1) % RB;
                if (nrb == 0) load_b(bfrag, buf, (ti + 1) / RB);
                acc[cur ^ 1] = f32x16{};
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    acc[cur ^ 1] = Ops::mfma(afrag[nrb], bfrag, s, acc[cur ^ 1]);
                    __builtin_amdgcn_sched_barrier(0);
                    epilogue(acc[cur], rb, (16 * s) / KS, (16 * (s + 1)) / KS, tile, c1, c2);
                    __builtin_amdgcn_sched_barrier(0);
                }
            } else {
                epilogue(acc[cur], rb, 0, 16, tile, c1, c2);
            }
            if (rb == RB - 1) {
                column_out(c1, c2, buf, cb);
                c1 = INFINITY;
                c2 = INFINITY;
            }
        }
        if (ch + 1 < nchunks) store(buf ^ 1);
        __syncthreads();
    }
    if (nchunks > 0) merge_cols(nchunks - 1);

    // Rows: resolve each lane's winning column, merge the 32 lanes of a half.
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            float k1 = m1[rb][r], k2 = m2[rb][r];
            // column of this lane's winner: tile id from the key, lane from l32
            uint32_t tl = __float_as_uint(-k1) & tmask;
            uint32_t j = k1 < INFINITY ? (tl * 32u + l32) : kNone;
            // the tile id wraps modulo 2^tbits only when tbits is too small; the
            // host sizes tbits so that it never does.
#pragma unroll
            for (int off = 1; off < 32; off <<= 1) {
                const float ok1 = __shfl_xor(k1, off);
                const float ok2 = __shfl_xor(k2, off);
                const uint32_t oj = __shfl_xor(j, off);
                const bool other = (ok1 < k1) || (ok1 == k1 && oj < j);
                const float n2 = other ? fminf(k1, ok2) : fminf(ok1, k2);
                if (other) {
                    k1 = ok1;
                    j = oj;
                }
                k2 = n2;
            }
            const uint32_t row =
                blk.row0 + wave * WR + rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (l32 == 0 && row < t.na) {
                const uint32_t g = t.row_base + row;
                out.row_k1[g] = k1;
                out.row_j1[g] = j;
                out.row_k2[g] = k2;
            }
        }
    }
}

// Per column of every pair: merge the workgroup partials.
__global__ void colmerge_kernel(const LowpTask* __restrict__ tasks, LowpOut out) {
    const LowpTask t = tasks[blockIdx.x];  // pairs on x: gridDim.y is capped at 65535
    const uint32_t nparts = (t.na + t.bm - 1) / t.bm;
    for (uint32_t col = blockIdx.y * blockDim.x + threadIdx.x; col < t.nb;
         col += gridDim.y * blockDim.x) {
        float a1 = INFINITY, a2 = INFINITY;
        for (uint32_t p = 0; p < nparts; ++p) {
            const float2 v = out.col_part[(size_t)t.part_base + (size_t)p * t.nb + col];
            const float n2 = fminf(fmaxf(a1, v.x), fminf(a2, v.y));
            a1 = fminf(a1, v.x);
            a2 = n2;
        }
        out.col_top2[t.col_base + col] = make_float2(a1, a2);
    }
}

__device__ __forceinline__ double pair_eps(const LowpTask& t, uint32_t tbits) {
    const ViewStats& A = *t.sa;
    const ViewStats& B = *t.sb;
    const double u = 0x1p-24;
    const double dot_q = (double)A.max_norm * B.max_res + (double)A.max_res * B.max_qnorm;
    const double qq = (double)A.max_qnorm * B.max_qnorm;
    const double nn = (double)A.max_sq + B.max_sq;
    if (t.fp8) {
        // |acc| bound: |q_a . q_b| + (n_a + n_b)/2 (the e4m3 hi+mid+lo folds are
        // within 2^-11 of n)
        const double acc = qq + 0.5 * nn * (1.0 + 0x1p-11);
        const double e = 2.0 * dot_q                  // e4m3 rounding of the components
                         + 2.0 * 2.0 * 200.0 * u * acc  // accumulation over K = 192 (2x margin)
                         + 2.0 * ldexp(1.0, (int)tbits - 22) * acc  // tile id in the low bits
                         + nn * 0x1p-11               // hi + mid + lo fold of the norms
                         + 64.0 * u * (nn + 2.0 * A.max_norm * (double)B.max_norm)
                         + 128.0 * u * nn;
        return e * 1.25 + 1e-6;
    }
    // |acc| bound: |q_a . q_b| + (n_a + n_b)/2 (the bf16 hi+lo norms are within 2^-15 of n)
    const double acc = qq + 0.5 * nn * (1.0 + 0x1p-15);
    double e = 2.0 * dot_q                      // bf16 rounding of the components
               + 2.0 * 2.0 * 132.0 * u * acc    // f32 accumulation inside the MFMA (2x margin)
               + 2.0 * ldexp(1.0, (int)tbits - 22) * acc  // tile id in the low bits
               + nn * 0x1p-15                   // hi + lo split of the norms
               + 64.0 * u * (nn + 2.0 * A.max_norm * (double)B.max_norm)  // reference f32 rounding
               + 128.0 * u * nn;                // f32 norms of the quantizer
    return e * 1.25 + 1e-6;
}

// eps[p] = the error bound of pair p (infinity when a view is scored exactly).
__global__ void pair_eps_kernel(const LowpTask* __restrict__ tasks, uint32_t n, uint32_t tbits,
                                double* __restrict__ eps) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const LowpTask t = tasks[p];
    eps[p] = (t.sa->bad | t.sb->bad) ? (double)INFINITY : pair_eps(t, tbits);
}

// Certain decisions; uncertain rows go to the exact queue.
// Verdict of row i from the approximate top-2 and the pair's error bound eps: verdict =
// the match or kNone; queue = the exact re-score decides; need_row = it needs the row scan
// (else only the column check of the certain j1).
__device__ __forceinline__ void decide_row(const LowpTask& t, const LowpOut& out, bool bad,
                                           double eps, float r2, int mutual, uint32_t i,
                                           uint32_t& verdict, bool& queue, bool& need_row,
                                           uint32_t& j1) {
    const uint32_t g = t.row_base + i;
    verdict = kNone;
    queue = false;
    need_row = true;
    j1 = kNone;
    if (t.nb < 2) {
        verdict = kNone;  // no second neighbour (matching.rs:206): never a match
    } else if (bad) {
        queue = true;
    } else {
        const double e1 = 2.0 * (double)out.row_k1[g];
        const double e2 = 2.0 * (double)out.row_k2[g];
        j1 = out.row_j1[g];
        const bool j1_certain = (e2 - e1) > 2.0 * eps && j1 < t.nb;
        const float lo1 = __double2float_rd(e1 - eps), hi1 = __double2float_ru(e1 + eps);
        const float lo2 = __double2float_rd(e2 - eps), hi2 = __double2float_ru(e2 + eps);
        const bool reject = lo1 >= r2 * hi2;
        const bool accept = hi1 < r2 * fmaxf(lo2, 0.0f);
        if (reject) {
            verdict = kNone;
        } else if (accept && j1_certain) {
            if (!mutual) {
                verdict = j1;
            } else {
                const float2 c = out.col_top2[t.col_base + j1];
                const double c1 = 2.0 * (double)c.x, c2 = 2.0 * (double)c.y;
                if (c2 - e1 > 2.0 * eps)
                    verdict = j1;
                else if (e1 - c1 > 2.0 * eps)
                    verdict = kNone;
                else {
                    queue = true;
                    need_row = false;
                }
            }
        } else {
            queue = true;
        }
    }
}

// One workgroup per pair. Pass 1 writes the certain verdicts and counts the queued rows; one
// reservation then gives the pair a contiguous run of queue entries, which pass 2 fills (the
// verdicts are recomputed, identically). The re-score groups consecutive entries, so a pair's
// queued rows share its streamed rows.
__global__ __launch_bounds__(256) void decide_kernel(const LowpTask* __restrict__ tasks,
                                                     LowpOut out, float ratio, int mutual,
                                                     uint32_t tbits) {
    __shared__ uint32_t wave_tot[4];
    __shared__ uint32_t run_base, chunk_base;
    const LowpTask t = tasks[blockIdx.x];  // one workgroup per pair
    const bool bad = (t.sa->bad | t.sb->bad) != 0;
    const double eps = pair_eps(t, tbits);
    const float r2 = ratio * ratio;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t mine = 0;
    for (uint32_t i = tid; i < t.na; i += blockDim.x) {
        uint32_t verdict, j1;
        bool queue, need_row;
        decide_row(t, out, bad, eps, r2, mutual, i, verdict, queue, need_row, j1);
        out.row_match[t.row_base + i] = verdict;
        mine += queue ? 1u : 0u;
    }
    // exclusive prefix of `mine` over the workgroup
    uint32_t incl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= (uint32_t)off) incl += o;
    }
    if (lane == 63) wave_tot[wave] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) {
        before += w < wave ? wave_tot[w] : 0u;
        total += wave_tot[w];
    }
    const uint32_t n_chunks =
        out.chunks && total >= kGatherMinRows ? (total + kChunkEntries - 1) / kChunkEntries : 0u;
    if (tid == 0 && total) {
        run_base = atomicAdd(&out.queue_len[0], total);
        if (n_chunks) chunk_base = atomicAdd(&out.queue_len[1], n_chunks);
    }
    __syncthreads();
    if (!total) return;
    for (uint32_t c = tid; c < n_chunks; c += blockDim.x)
        out.chunks[chunk_base + c] =
            make_uint2(run_base + c * kChunkEntries, min(kChunkEntries, total - c * kChunkEntries));
    uint32_t slot = run_base + before + incl - mine;
    for (uint32_t i = tid; i < t.na; i += blockDim.x) {
        uint32_t verdict, j1;
        bool queue, need_row;
        decide_row(t, out, bad, eps, r2, mutual, i, verdict, queue, need_row, j1);
        if (queue) {
            if (out.chunks) out.pre[slot].w = 0u;
            out.queue[slot++] = make_uint4(blockIdx.x, i, need_row ? kNone : j1, 0u);
        }
    }
}

// Gathered exact row scans: the row-scan entries of a chunk (up to 64 queue
// entries of one pair) holding at least kGatherMinRows of them, their a rows
// staged transposed, against all of b with the exact FP32 sweep (the
// reference's formula and row update order); results per queue entry in pre.
// Sparser chunks leave their rows to rescore_kernel, whose per-row scans cost
// less when few rows share b.
__global__ __launch_bounds__(kExactNT) void exact_rows_kernel(const LowpTask* __restrict__ tasks,
                                                              LowpOut out) {
    __shared__ float As[kDescLen][kExactBM];
    __shared__ float Bs[kDescLen][kExactBN];
    __shared__ uint32_t slot_of[kExactBM];
    __shared__ uint32_t n_rows;
    const uint32_t n_chunks = out.queue_len[1];
    const int tid = threadIdx.x;
    for (uint32_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        const uint2 ck = out.chunks[c];
        const uint32_t k0 = ck.x;
        __syncthreads();  // the previous chunk's readers of As / slot_of are done
        if (tid < 64) {  // one wave compacts the chunk's row-scan entries
            bool scan = false;
            if ((uint32_t)tid < ck.y) {
                const uint4 e = out.queue[k0 + tid];
                scan = e.z == kNone && e.w == 0u;  // w != 0: settled by the BF16 re-rank
            }
            const unsigned long long m = __ballot(scan);
            if (scan) slot_of[__popcll(m & ((1ull << tid) - 1ull))] = k0 + tid;
            if (tid == 0) n_rows = __popcll(m);
        }
        __syncthreads();
        const uint32_t rows = n_rows;
        if (rows < kGatherMinRows) continue;  // uniform: n_rows is read after the barrier
        const LowpTask t = tasks[out.queue[k0].x];
        for (int e = tid; e < kExactBM * (kDescLen / 4); e += kExactNT) {
            const int r = e / (kDescLen / 4), c4 = e % (kDescLen / 4);
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if ((uint32_t)r < rows)
                v = *reinterpret_cast<const float4*>(
                    t.a + (size_t)out.queue[slot_of[r]].y * kDescLen + 4 * c4);
            As[4 * c4 + 0][r] = v.x;
            As[4 * c4 + 1][r] = v.y;
            As[4 * c4 + 2][r] = v.z;
            As[4 * c4 + 3][r] = v.w;
        }
        Top2 st[4];
        exact_sweep(As, Bs, t.b, t.nb, st, [](const float (&)[4][4], uint32_t, uint32_t) {});
        const int tx = tid & 15, ty = tid >> 4;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const Top2 m = exact_row_merge(st[r]);
            const uint32_t rr = ty * 4 + r;
            if (tx == 0 && rr < rows)
                out.pre[slot_of[rr]] = make_uint4(m.j, __float_as_uint(m.d1), __float_as_uint(m.d2), 1u);
        }
    }
}

struct Top2x {
    uint32_t j;
    float d1, d2;
};

__device__ __forceinline__ Top2x merge2(Top2x a, Top2x b) {
    Top2x r;
    const bool bw = (b.d1 < a.d1) || (b.d1 == a.d1 && b.j < a.j);
    if (bw) {
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

// Exact argmin of column j over all rows of a (lowest i among equal d;
// d = inf and NaN never win, matching.rs:188-203). Returns kNone if none.
constexpr int kRescoreGroup = 8;  // queued rows of one pair re-scored together

// Exact squared distances, the reference's formula (matching.rs:146-155: eight running sums,
// lane k taking components k, k + 8, ..., no FMA, then the eight lanes added in order), from
// the streamed row x to the first m rows of ys (LDS): d[r] = |x - ys[r]|^2. A difference and
// its negation square to the same value, so x may be either side.
__device__ __forceinline__ void exact_dist_group(const float* __restrict__ x,
                                                 const float (*ys)[kDescLen], int m,
                                                 float* d, uint32_t need = 0xffu) {
    float acc[kRescoreGroup][8];
#pragma unroll
    for (int r = 0; r < kRescoreGroup; ++r)
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[r][k] = 0.f;
#pragma unroll 2
    for (int c = 0; c < kDescLen; c += 8) {
        const float4 x0 = *reinterpret_cast<const float4*>(x + c);
        const float4 x1 = *reinterpret_cast<const float4*>(x + c + 4);
#pragma unroll
        for (int r = 0; r < kRescoreGroup; ++r) {
            if (r < m && ((need >> r) & 1u)) {
                const float4 y0 = *reinterpret_cast<const float4*>(&ys[r][c]);
                const float4 y1 = *reinterpret_cast<const float4*>(&ys[r][c + 4]);
                float e;
                e = x0.x - y0.x; acc[r][0] = acc[r][0] + e * e;
                e = x0.y - y0.y; acc[r][1] = acc[r][1] + e * e;
                e = x0.z - y0.z; acc[r][2] = acc[r][2] + e * e;
                e = x0.w - y0.w; acc[r][3] = acc[r][3] + e * e;
                e = x1.x - y1.x; acc[r][4] = acc[r][4] + e * e;
                e = x1.y - y1.y; acc[r][5] = acc[r][5] + e * e;
                e = x1.z - y1.z; acc[r][6] = acc[r][6] + e * e;
                e = x1.w - y1.w; acc[r][7] = acc[r][7] + e * e;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kRescoreGroup; ++r) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) s = s + acc[r][k];
        d[r] = s;
    }
}

// Exact row top-2 (matching.rs:182-187) of the m rows in ys against all of b: thread k
// takes columns k, k + 256, ... in order, then the 256 states merge (lanes, then waves).
__device__ void exact_row_top2_group(const LowpTask& t, const float (*ys)[kDescLen], int m,
                                     Top2x (*red)[4], Top2x* res) {
    Top2x st[kRescoreGroup];
#pragma unroll
    for (int r = 0; r < kRescoreGroup; ++r) st[r] = Top2x{kNone, INFINITY, INFINITY};
    for (uint32_t j = threadIdx.x; j < t.nb; j += blockDim.x) {
        float d[kRescoreGroup];
        exact_dist_group(t.b + (size_t)j * kDescLen, ys, m, d);
#pragma unroll
        for (int r = 0; r < kRescoreGroup; ++r) {
            if (d[r] < st[r].d1) {
                st[r].d2 = st[r].d1;
                st[r].d1 = d[r];
                st[r].j = j;
            } else if (d[r] < st[r].d2) {
                st[r].d2 = d[r];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kRescoreGroup; ++r) {
        if (r >= m) break;
        Top2x x = st[r];
        for (int off = 32; off >= 1; off >>= 1) {
            Top2x o;
            o.j = __shfl_xor(x.j, off);
            o.d1 = __shfl_xor(x.d1, off);
            o.d2 = __shfl_xor(x.d2, off);
            x = merge2(x, o);
        }
        if ((threadIdx.x & 63) == 0) red[r][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    for (int r = 0; r < m; ++r) {
        Top2x x = red[r][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) x = merge2(x, red[r][w]);
        res[r] = x;
    }
    __syncthreads();
}

// Exact column argmin (lowest row index among equal distances) of the m b rows in ys over
// a; kNone when every distance is infinite. Only the row blocks that can hold the argmin are
// scanned: with eps bounding |2 key - d^2| for every element of the pair, a row r can be the
// exact minimum (or tie with it) of column j only if 2 key(r, j) <= 2 C1(j) + 2 eps, C1 the
// column's smallest key, so a row block whose own smallest key for j (its column partial)
// exceeds that bound holds no candidate for j; every row of an unscanned block is strictly
// farther than the exact minimum. The blocks scanned are the union over the m columns
// (scanning extra rows never changes an exact argmin). cols[r] is the column index of ys[r].
__device__ void exact_col_argmin_group(const LowpTask& t, const LowpOut& out, double eps,
                                       const uint32_t* cols, const float (*ys)[kDescLen], int m,
                                       unsigned long long (*red)[4],
                                       unsigned long long (*blk_mask)[4], uint32_t* res) {
    const uint32_t bm = t.bm ? t.bm : 128u;
    const uint32_t nblk = (t.na + bm - 1) / bm;
    // candidate row blocks of each column (all of them when the block size is unknown,
    // there are more than 64 blocks or eps is not finite)
    const bool all = t.bm == 0 || nblk > 64 || !(eps < INFINITY);
    for (int r = 0; r < m; ++r) {
        bool cand = false;
        const uint32_t b = threadIdx.x;
        if (!all && b < nblk) {
            const uint32_t j = cols[r];
            const double thr = 2.0 * (double)out.col_top2[t.col_base + j].x + 2.0 * eps;
            const double kb = 2.0 * (double)out.col_part[t.part_base + (size_t)b * t.nb + j].x;
            cand = kb <= thr;
        }
        const unsigned long long bits = __ballot(cand);
        if ((threadIdx.x & 63) == 0) blk_mask[r][threadIdx.x >> 6] = bits;
    }
    __syncthreads();
    unsigned long long mask[kRescoreGroup];
#pragma unroll
    for (int r = 0; r < kRescoreGroup; ++r)
        mask[r] = r >= m ? 0ull
                         : (all ? ~0ull
                                : (blk_mask[r][0] | blk_mask[r][1] | blk_mask[r][2] | blk_mask[r][3]));
    unsigned long long best[kRescoreGroup];
#pragma unroll
    for (int r = 0; r < kRescoreGroup; ++r) best[r] = kColEmpty;
    for (uint32_t b = 0; b < nblk; ++b) {
        // the columns that need this block (uniform across the workgroup)
        uint32_t need = 0;
#pragma unroll
        for (int r = 0; r < kRescoreGroup; ++r) need |= (uint32_t)((mask[r] >> (b & 63u)) & 1ull) << r;
        if (!need) continue;
        const uint32_t i_end = min(t.na, (b + 1) * bm);
        for (uint32_t i = b * bm + threadIdx.x; i < i_end; i += blockDim.x) {
            float d[kRescoreGroup];
            exact_dist_group(t.a + (size_t)i * kDescLen, ys, m, d, need);
#pragma unroll
            for (int r = 0; r < kRescoreGroup; ++r) {
                if (((need >> r) & 1u) && d[r] < INFINITY) {
                    const unsigned long long key =
                        ((unsigned long long)__float_as_uint(d[r]) << 32) | i;
                    best[r] = key < best[r] ? key : best[r];
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kRescoreGroup; ++r) {
        if (r >= m) break;
        unsigned long long b = best[r];
        for (int off = 32; off >= 1; off >>= 1) {
            const unsigned long long o = __shfl_xor(b, off);
            b = o < b ? o : b;
        }
        if ((threadIdx.x & 63) == 0) red[r][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    for (int r = 0; r < m; ++r) {
        unsigned long long b = red[r][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = red[r][w] < b ? red[r][w] : b;
        res[r] = b == kColEmpty ? kNone : (uint32_t)(b & 0xffffffffu);
    }
    __syncthreads();
}

// BF16 re-rank of the FP8 path's open row decisions. The FP8 ranking's error
// bound is about 15 times BF16's, so it leaves ~10x more rows undecided; each
// would cost an exact FP32 scan of all of b. Instead, per chunk (up to 64 queue
// entries of one pair), the entries that still need a row scan are ranked again
// against all of b with the BF16 MFMA form of mfma_rank_kernel (K = 144 with
// the norm fold, four waves = 2 row tiles x 2 column tiles of each 64-column
// chunk, row top-2 only) and decided with the BF16 pair bound eps_b:
//   certain reject (lo1 >= r^2 hi2)                          -> entry.w = 1;
//   certain accept with j1 certain, not mutual                -> z = j1, w = 2;
//   mutual: the FP8 column state (c1, c2 of j1, bound eps_8) decides it when
//     2 c2 - eps_8 > e1 + eps_b (row i is the column's unique exact winner: any
//     other row's exact distance is at least 2 c2 - eps_8, except the row
//     holding c1, which must then be i) -> z = j1, w = 2, or when
//     e1 - eps_b > 2 c1 + eps_8 (a closer row exists)      -> w = 1;
//     otherwise z = j1, w = 0: the exact column check of rescore_kernel;
//   anything else stays a row scan (exact_rows_kernel / rescore_kernel).
// COL = true is the column form for the open mutual checks (entries with j1
// known, w = 0): the gathered vectors are the b rows j1, swept against all of
// a, which gives each column's BF16 top-2 over the rows with the argmin row r1.
// When k2 - k1 > eps_b, r1 is the column's unique exact winner (its exact
// distance is at most 2 k1 + eps_b, every other row's at least 2 k2 - eps_b):
// the entry is accepted (w = 2) if r1 is its row, else rejected (w = 1).
// Every decision follows from bounds that hold for every element of the pair,
// so the match list is the exact FP32 one whatever this kernel settles.
template <bool COL>
__global__ __launch_bounds__(256) void rerank_bf16_kernel(const LowpTask* __restrict__ tasks_b,
                                                          const LowpTask* __restrict__ tasks_8,
                                                          LowpOut out, float ratio, int mutual,
                                                          uint32_t tbits_8, uint32_t tbits_g) {
    using Ops = RankOps<false>;
    constexpr int STRIDE = Ops::kStride;
    constexpr int PIECES = Ops::kRowBytes / 16;  // 16 per bf16 row
    __shared__ __attribute__((aligned(16))) uint8_t Bs[64 * STRIDE];
    __shared__ uint32_t Bn[64];
    __shared__ uint32_t slot_of[64];
    __shared__ uint32_t n_rows;
    __shared__ float mk1[2][32], mk2[2][32];  // the column-tile-1 waves' row results
    __shared__ uint32_t mj[2][32];
    const uint32_t n_chunks = out.queue_len[1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l32 = lane & 31;
    const bool lo_half = lane < 32;
    const int rt = wave >> 1, ct = wave & 1;  // this wave's row tile and column tile
    const uint32_t tmask = (1u << tbits_g) - 1u;
    const float r2 = ratio * ratio;
    for (uint32_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        const uint2 ck = out.chunks[c];
        const uint32_t k0 = ck.x;
        __syncthreads();  // the previous chunk's readers of slot_of / Bs are done
        if (tid < 64) {
            bool scan = false;
            if ((uint32_t)tid < ck.y) {
                const uint4 e = out.queue[k0 + tid];
                scan = e.w == 0u && (COL ? e.z != kNone : e.z == kNone);
            }
            const unsigned long long m = __ballot(scan);
            if (scan) slot_of[__popcll(m & ((1ull << tid) - 1ull))] = k0 + tid;
            if (tid == 0) n_rows = __popcll(m);
        }
        __syncthreads();
        const uint32_t rows = n_rows;
        if (rows == 0) continue;  // uniform
        const uint32_t pair = out.queue[k0].x;
        const LowpTask t = tasks_b[pair];
        const LowpTask t8 = tasks_8[pair];
        if ((t.sa->bad | t.sb->bad | t8.sa->bad | t8.sb->bad) != 0) continue;  // exact path
        // the gathered side (rows of a; for COL the b rows j1) and the swept side
        const uint8_t* gsrc = COL ? Ops::rows_b(t) : Ops::rows_a(t);
        const uint8_t* qb = COL ? Ops::rows_a(t) : Ops::rows_b(t);
        const uint32_t nsw = COL ? t.na : t.nb;
        // a fragments of this wave's row tile (gathered vectors, padding beyond `rows`)
        const uint32_t lr = (uint32_t)(rt * 32 + l32);
        const bool valid = lr < rows;
        uint32_t arow = 0u;
        if (valid) {
            const uint4 e = out.queue[slot_of[lr]];
            arow = COL ? e.z : e.y;
        }
        typename Ops::AFrag afrag;
        Ops::load_a(afrag, gsrc + (size_t)arow * Ops::kRowBytes, valid,
                    valid ? (COL ? Ops::meta_b(t, arow) : Ops::meta_a(t, arow)) : Ops::pad_meta(),
                    lo_half);
        float m1[16], m2[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            m1[r] = INFINITY;
            m2[r] = INFINITY;
        }
        uint32_t keep_v;
        asm volatile("v_mov_b32 %0, %1" : "=v"(keep_v) : "s"(~tmask));
        for (uint32_t col0 = 0; col0 < nsw; col0 += 64) {
            __syncthreads();  // the previous chunk's fragments are read
#pragma unroll
            for (int k = 0; k < 64 * PIECES / 256; ++k) {
                const int e = tid + 256 * k, cc = e / PIECES, seg = e % PIECES;
                const uint32_t col = col0 + cc;
                const uint4 v = col < nsw ? ldg_u4(qb + (size_t)col * Ops::kRowBytes + 16 * seg)
                                          : make_uint4(0, 0, 0, 0);
                *reinterpret_cast<uint4*>(&Bs[cc * STRIDE + 16 * seg]) = v;
            }
            if (tid < 64)
                Bn[tid] = col0 + tid >= nsw ? Ops::pad_meta()
                          : COL             ? Ops::meta_a(t, col0 + tid)
                                            : Ops::meta_b(t, col0 + tid);
            __syncthreads();
            typename Ops::BFrag bfrag;
            Ops::load_b(bfrag, &Bs[(ct * 32 + l32) * STRIDE], Bn[ct * 32 + l32], lo_half);
            f32x16 acc = f32x16{};
#pragma unroll
            for (int st = 0; st < Ops::kSteps; ++st) acc = Ops::mfma(afrag, bfrag, st, acc);
            const uint32_t tile = ((col0 >> 5) + (uint32_t)ct) & tmask;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float pk = __uint_as_float((__float_as_uint(acc[r]) & keep_v) | tile);
                top2_neg(m1[r], m2[r], pk);
            }
        }
        // resolve each register's row over the 32 lanes of a half, then the two
        // column tiles (waves ct = 0, 1 of the same row tile)
        float rk1[16], rk2[16];
        uint32_t rj[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            float k1 = m1[r], k2 = m2[r];
            const uint32_t tl = __float_as_uint(-k1) & tmask;
            uint32_t j = k1 < INFINITY ? (tl * 32u + l32) : kNone;
#pragma unroll
            for (int off = 1; off < 32; off <<= 1) {
                const float ok1 = __shfl_xor(k1, off);
                const float ok2 = __shfl_xor(k2, off);
                const uint32_t oj = __shfl_xor(j, off);
                const bool other = (ok1 < k1) || (ok1 == k1 && oj < j);
                const float n2 = other ? fminf(k1, ok2) : fminf(ok1, k2);
                if (other) {
                    k1 = ok1;
                    j = oj;
                }
                k2 = n2;
            }
            rk1[r] = k1;
            rk2[r] = k2;
            rj[r] = j;
        }
        __syncthreads();
        if (ct == 1 && l32 == 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                mk1[rt][row] = rk1[r];
                mk2[rt][row] = rk2[r];
                mj[rt][row] = rj[r];
            }
        }
        __syncthreads();
        if (ct == 0 && l32 == 0) {
            const double eps_b = pair_eps(t, tbits_g);
            const double eps_8 = pair_eps(t8, tbits_8);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const uint32_t g = (uint32_t)(rt * 32 + row);
                if (g >= rows) continue;
                // merge with the other column tile (same rule as the lane merge)
                float k1 = rk1[r], k2 = rk2[r];
                uint32_t j = rj[r];
                const float ok1 = mk1[rt][row], ok2 = mk2[rt][row];
                const uint32_t oj = mj[rt][row];
                const bool other = (ok1 < k1) || (ok1 == k1 && oj < j);
                const float n2 = other ? fminf(k1, ok2) : fminf(ok1, k2);
                if (other) {
                    k1 = ok1;
                    j = oj;
                }
                k2 = n2;
                const double e1 = 2.0 * (double)k1, e2 = 2.0 * (double)k2;
                const uint32_t slot = slot_of[g];
                if (COL) {
                    // j: the column's BF16 argmin row, its exact winner when the gap exceeds eps_b
                    if (j < t.na && (e2 - eps_b) > (e1 + eps_b))
                        out.queue[slot].w = j == out.queue[slot].y ? 2u : 1u;
                    continue;
                }
                const bool j1_certain = (e2 - e1) > 2.0 * eps_b && j < t.nb;
                const float lo1 = __double2float_rd(e1 - eps_b), hi1 = __double2float_ru(e1 + eps_b);
                const float lo2 = __double2float_rd(e2 - eps_b), hi2 = __double2float_ru(e2 + eps_b);
                if (lo1 >= r2 * hi2) {
                    out.queue[slot].w = 1u;  // certain reject
                } else if (hi1 < r2 * fmaxf(lo2, 0.0f) && j1_certain) {
                    uint32_t w = 2u;
                    if (mutual) {
                        const float2 cc = out.col_top2[t8.col_base + j];
                        const double c1 = 2.0 * (double)cc.x, c2 = 2.0 * (double)cc.y;
                        if (c2 - eps_8 > e1 + eps_b)
                            w = 2u;
                        else if (e1 - eps_b > c1 + eps_8)
                            w = 1u;
                        else
                            w = 0u;  // the exact column check decides
                    }
                    out.queue[slot].z = j;
                    out.queue[slot].w = w;
                }
            }
        }
    }
}

// Queued rows, kRescoreGroup consecutive queue entries per step; consecutive entries of one
// pair (the decide kernel reserves a wave's entries together) share each streamed row of b
// (row scans) or of a (column scans).
__global__ __launch_bounds__(256) void rescore_kernel(const LowpTask* __restrict__ tasks,
                                                      LowpOut out, float ratio, int mutual,
                                                      uint32_t tbits) {
    __shared__ __attribute__((aligned(16))) float ys[kRescoreGroup][kDescLen];
    __shared__ Top2x red_t[kRescoreGroup][4];
    __shared__ unsigned long long red_c[kRescoreGroup][4];
    __shared__ unsigned long long blk_mask[kRescoreGroup][4];
    __shared__ Top2x row_res[kRescoreGroup];
    __shared__ uint32_t col_res[kRescoreGroup];
    const float r2 = ratio * ratio;
    const uint32_t n = out.queue_len[0];
    for (uint32_t base = blockIdx.x * kRescoreGroup; base < n;
         base += gridDim.x * kRescoreGroup) {
        const uint32_t kend = min(base + (uint32_t)kRescoreGroup, n);
        uint32_t k0 = base;
        while (k0 < kend) {
            // the run [k0, k1) of entries of one pair
            const uint32_t pair = out.queue[k0].x;
            uint32_t k1 = k0 + 1;
            while (k1 < kend && out.queue[k1].x == pair) ++k1;
            const LowpTask t = tasks[pair];
            const bool bad = (t.sa->bad | t.sb->bad) != 0;
            const double eps = pair_eps(t, tbits);
            const int m = (int)(k1 - k0);
            // Entry classes as bit masks over r (bit r: entry k0 + r): j1 known (ratio
            // accepted with certainty), row scan done by exact_rows_kernel, row scan here.
            // A row scan's slot in ys / row_res is the count of row scans below it.
            uint32_t row_i[kRescoreGroup], j1[kRescoreGroup];
            bool accept[kRescoreGroup], col_needed[kRescoreGroup];
            uint32_t known_m = 0, pre_m = 0, scan_m = 0, done_m = 0;
#pragma unroll
            for (int r = 0; r < kRescoreGroup; ++r) {
                row_i[r] = kNone;
                j1[r] = kNone;
                bool settled_accept = false;
                if (r < m) {
                    const uint4 item = out.queue[k0 + r];
                    row_i[r] = item.y;
                    j1[r] = item.z;
                    if (item.w != 0u) {  // settled by the BF16 re-rank: 1 reject, 2 accept
                        done_m |= 1u << r;
                        settled_accept = item.w == 2u;
                    } else if (item.z != kNone) known_m |= 1u << r;
                    else if (out.chunks && out.pre[k0 + r].w) pre_m |= 1u << r;
                    else scan_m |= 1u << r;
                }
                accept[r] = ((known_m >> r) & 1u) || settled_accept;
                col_needed[r] = r < m && !((done_m >> r) & 1u);
            }
            const int mr = __popc(scan_m);
            // row scans of the entries that need one
            if (mr > 0) {
                __syncthreads();
#pragma unroll
                for (int r = 0; r < kRescoreGroup; ++r)
                    if (((scan_m >> r) & 1u) && threadIdx.x < kDescLen)
                        ys[__popc(scan_m & ((1u << r) - 1u))][threadIdx.x] =
                            t.a[(size_t)row_i[r] * kDescLen + threadIdx.x];
                __syncthreads();
                exact_row_top2_group(t, ys, mr, red_t, row_res);
            }
#pragma unroll
            for (int r = 0; r < kRescoreGroup; ++r) {
                if (!(((scan_m | pre_m) >> r) & 1u)) continue;
                Top2x st;
                if ((pre_m >> r) & 1u) {
                    const uint4 pr = out.pre[k0 + r];
                    st = Top2x{pr.x, __uint_as_float(pr.y), __uint_as_float(pr.z)};
                } else {
                    st = row_res[__popc(scan_m & ((1u << r) - 1u))];
                }
                j1[r] = st.j;
                accept[r] = st.j != kNone && isfinite(st.d2) && !(st.d1 >= r2 * st.d2);
                if (accept[r] && mutual && !bad) {
                    const float2 c = out.col_top2[t.col_base + st.j];
                    const double d1 = st.d1;
                    if ((2.0 * (double)c.y) - d1 > eps) {
                        col_needed[r] = false;  // row i is the unique column winner
                    } else if (d1 - 2.0 * (double)c.x > eps) {
                        accept[r] = false;  // another row is strictly closer
                        col_needed[r] = false;
                    }
                }
            }
            // column scans of the accepted entries whose mutual check is open (one per
            // distinct j1)
            int col_slot[kRescoreGroup];
            bool col_first[kRescoreGroup];
            int mc = 0;
#pragma unroll
            for (int r = 0; r < kRescoreGroup; ++r) {
                col_slot[r] = -1;
                col_first[r] = false;
                if (r < m && accept[r] && mutual && col_needed[r]) {
#pragma unroll
                    for (int q = 0; q < r; ++q)
                        if (col_slot[r] < 0 && col_slot[q] >= 0 && j1[q] == j1[r]) col_slot[r] = col_slot[q];
                    if (col_slot[r] < 0) {
                        col_slot[r] = mc++;
                        col_first[r] = true;
                    }
                }
            }
            if (mc > 0) {
                uint32_t col_j[kRescoreGroup];
                __syncthreads();
#pragma unroll
                for (int r = 0; r < kRescoreGroup; ++r) {
                    if (col_first[r]) col_j[col_slot[r]] = j1[r];
                    if (col_first[r] && threadIdx.x < kDescLen)
                        ys[col_slot[r]][threadIdx.x] = t.b[(size_t)j1[r] * kDescLen + threadIdx.x];
                }
                __syncthreads();
                exact_col_argmin_group(t, out, bad ? (double)INFINITY : eps, col_j, ys, mc, red_c,
                                       blk_mask, col_res);
#pragma unroll
                for (int r = 0; r < kRescoreGroup; ++r)
                    if (col_slot[r] >= 0) accept[r] = col_res[col_slot[r]] == row_i[r];
            }
            if (threadIdx.x == 0) {
#pragma unroll
                for (int r = 0; r < kRescoreGroup; ++r)
                    if (r < m) out.row_match[t.row_base + row_i[r]] = accept[r] ? j1[r] : kNone;
            }
            k0 = k1;
        }
    }
}

}  // namespace

void launch_stats_reset(ViewStats* st, uint32_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(stats_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, s, st, n);
}

void launch_quantize(const float* x, uint32_t n, uint16_t* q, uint32_t* nrm, ViewStats* st,
                     hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(quantize_kernel, dim3((n + 255) / 256), dim3(256), 0, s, x, n, q, nrm, st);
}

void launch_quantize_views(const float* x, const uint64_t* d_off, uint32_t n_views,
                           uint32_t max_rows, uint16_t* q, uint32_t* nrm, ViewStats* st,
                           hipStream_t s) {
    if (!n_views || !max_rows) return;
    hipLaunchKernelGGL(quantize_views_kernel, dim3(n_views, (max_rows + 255) / 256), dim3(256), 0,
                       s, x, d_off, q, nrm, st);
}

void launch_quantize_fp8(const float* x, uint32_t n, uint8_t* q, Fp8Meta* meta, ViewStats* st,
                        hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(quantize_fp8_kernel, dim3((n + 255) / 256), dim3(256), 0, s, x, n, q, meta,
                       st);
}

void launch_quantize_views_fp8(const float* x, const uint64_t* d_off, uint32_t n_views,
                               uint32_t max_rows, uint8_t* q, Fp8Meta* meta, ViewStats* st,
                               hipStream_t s) {
    if (!n_views || !max_rows) return;
    hipLaunchKernelGGL(quantize_views_fp8_kernel, dim3(n_views, (max_rows + 255) / 256), dim3(256),
                       0, s, x, d_off, q, meta, st);
}

uint32_t lowp_block_rows(int tile) { return tile_rb(tile) * 128u; }

template <bool F8>
static void launch_rank(int tile, const LowpTask* tasks, const RowBlock* blocks, uint32_t n_blocks,
                        const LowpOut& out, uint32_t tbits, hipStream_t s) {
    switch (tile) {
        case kTile128x64:
            hipLaunchKernelGGL((mfma_rank_kernel<F8, 1, 2>), dim3(n_blocks), dim3(256), 0, s, tasks,
                               blocks, out, tbits);
            break;
        case kTile128x128:
            hipLaunchKernelGGL((mfma_rank_kernel<F8, 1, 4>), dim3(n_blocks), dim3(256), 0, s, tasks,
                               blocks, out, tbits);
            break;
        case kTile256x64:
            hipLaunchKernelGGL((mfma_rank_kernel<F8, 2, 2>), dim3(n_blocks), dim3(256), 0, s, tasks,
                               blocks, out, tbits);
            break;
        default:
            hipLaunchKernelGGL((mfma_rank_kernel<F8, 2, 4>), dim3(n_blocks), dim3(256), 0, s, tasks,
                               blocks, out, tbits);
            break;
    }
}

void launch_mfma_rank(int tile, bool fp8, const LowpTask* tasks, const RowBlock* blocks,
                      uint32_t n_blocks, const LowpOut& out, uint32_t tbits, hipStream_t s) {
    if (!n_blocks) return;
    if (fp8)
        launch_rank<true>(tile, tasks, blocks, n_blocks, out, tbits, s);
    else
        launch_rank<false>(tile, tasks, blocks, n_blocks, out, tbits, s);
}

void launch_pair_eps(const LowpTask* tasks, uint32_t n_pairs, uint32_t tbits, double* eps,
                     hipStream_t s) {
    if (!n_pairs) return;
    hipLaunchKernelGGL(pair_eps_kernel, dim3((n_pairs + 255) / 256), dim3(256), 0, s, tasks,
                       n_pairs, tbits, eps);
}

void launch_colmerge(const LowpTask* tasks, uint32_t n_pairs, const LowpOut& out, hipStream_t s) {
    if (!n_pairs) return;
    hipLaunchKernelGGL(colmerge_kernel, dim3(n_pairs, 8), dim3(256), 0, s, tasks, out);
}

void launch_decide(const LowpTask* tasks, uint32_t n_pairs, const LowpOut& out, float ratio,
                   int mutual, uint32_t tbits, hipStream_t s) {
    if (!n_pairs) return;
    hipLaunchKernelGGL(decide_kernel, dim3(n_pairs), dim3(256), 0, s, tasks, out, ratio, mutual,
                       tbits);
}

void launch_rerank_bf16(const LowpTask* tasks_b, const LowpTask* tasks_8, const LowpOut& out,
                        float ratio, int mutual, uint32_t tbits_8, uint32_t tbits_g,
                        uint32_t grid, hipStream_t s) {
    if (!out.chunks) return;
    hipLaunchKernelGGL(rerank_bf16_kernel<false>, dim3(std::max(grid, 1u)), dim3(256), 0, s,
                       tasks_b, tasks_8, out, ratio, mutual, tbits_8, tbits_g);
    if (mutual)  // the open column checks, including the ones the row pass just opened
        hipLaunchKernelGGL(rerank_bf16_kernel<true>, dim3(std::max(grid, 1u)), dim3(256), 0, s,
                           tasks_b, tasks_8, out, ratio, mutual, tbits_8, tbits_g);
}

void launch_rescore(const LowpTask* tasks, const LowpOut& out, float ratio, int mutual,
                    uint32_t tbits, uint32_t grid, hipStream_t s) {
    if (out.chunks)
        hipLaunchKernelGGL(exact_rows_kernel, dim3(std::max(grid / 2, 1u)), dim3(kExactNT), 0, s,
                           tasks, out);
    hipLaunchKernelGGL(rescore_kernel, dim3(grid), dim3(256), 0, s, tasks, out, ratio, mutual,
                       tbits);
}

}  // namespace skh
