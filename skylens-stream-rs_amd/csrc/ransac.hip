// ransac_fundamental on gfx950.
//
// Reference: crates/core/src/matching.rs:941-1104 (and the helpers it calls).
// The sequential semantics (adaptive iteration count, local-optimisation
// trigger, 5-entry Jaccard pool, stable sort) are replayed exactly while the
// heavy work runs across lanes:
//   * a persistent launch of four-wave workgroups takes pairs from a queue in
//     descending order of a first-batch cost estimate; the costliest pairs run
//     on a whole workgroup, the others on one wave each;
//   * hypotheses in batches of 63: the lead wave draws the seeded samples
//     (matching.rs:956-979), solver groups of three lanes solve one 8-point F
//     each (FP64) and count its Sampson inliers;
//   * local optimisation for hypothesis k depends only on F_k and the data
//     (matching.rs:993-1015), so every hypothesis that *could* trigger it runs
//     it: cnt >= 8 and 4 cnt >= max(best count at batch start, raw counts
//     >= 8 earlier in the batch), a superset of the real triggers because the
//     best count is never below an earlier raw count; point sums run
//     wave-cooperatively in the host's order, the 8-point solves of all
//     candidates (and of the next batch's hypotheses) share solver rounds;
//   * the lead wave replays the batch in order with the precomputed results;
//     pool Jaccard tests use inlier bitsets;
//   * the pool's Sampson-LM trajectories (matching.rs:1041-1057), the final
//     refits, gates and GRIC model selection follow the reference.
// All FP64 routines match the host restatement operation for operation.
#pragma clang fp contract(off)

#include "ransac_dev.hpp"
#include "skh_ransac.hpp"

#include <algorithm>

namespace skh {
namespace {

using namespace rdev;

constexpr int NL = 64;     // lanes per wave
constexpr int kWaves = 4;  // waves per workgroup: four small pairs, or one large pair

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x >> 6); }

__device__ __forceinline__ double readlane_d(double v, int l) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Orders this wave's LDS and device-memory accesses before the following
// ones: the lanes of a wave exchange data through LDS and device scratch,
// and the waves of a workgroup may work on different pairs.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Orders this wave's LDS accesses before the following ones (a staged
// chunk's LDS writes before the lanes read each other's entries) without
// waiting for its outstanding device-memory loads: LDS operations of one wave
// are performed in program order, so this only waits for the wave's own LDS
// operations to finish and keeps the compiler from moving LDS accesses across.
__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt not waited on
    __builtin_amdgcn_wave_barrier();
}
// Loads from device (global) memory through the global address space: the
// compiler then issues global loads, which only the vector-memory counter
// tracks, instead of flat loads, which the LDS counter tracks too and which
// force full waits before every LDS access that follows them.
__device__ __forceinline__ double2 ldg2(const double2* p, uint32_t i) {
    const __attribute__((address_space(1))) double* g =
        (const __attribute__((address_space(1))) double*)p;
    return make_double2(g[2 * (size_t)i], g[2 * (size_t)i + 1]);
}
__device__ __forceinline__ uint32_t ldg(const uint32_t* p, uint32_t i) {
    return ((const __attribute__((address_space(1))) uint32_t*)p)[i];
}
__device__ __forceinline__ void load9(const double* src, double* dst) {
#pragma unroll
    for (int i = 0; i < 9; ++i) dst[i] = src[i];
}

__constant__ double kNM[6] = {3.0, 2.0, 1.5, 1.0, 1.0, 1.0};

__host__ __device__ __forceinline__ uint32_t words(uint32_t n) { return (n + 31) / 32; }

// The first four of eight gathered samples (a homography hypothesis).
struct Pts4 {
    const Pts8* b;
    static constexpr uint32_t n = 4;
    __device__ bool take(uint32_t) const { return true; }
    __device__ double2 p(uint32_t i) const { return b->a[i]; }
    __device__ double2 q(uint32_t i) const { return b->b[i]; }
};


// ---------------------------------------------------------------- Cholesky

// Lower triangle of a symmetric 9x9 in ut() slots; solves m x = rhs by its
// LDL^T factorisation (the host's cholesky_solve9 step for step: right-
// looking column updates, unit lower L, diagonal D and its reciprocals, then
// L y = rhs, y *= D^-1, L^T x = y). Fully expanded at compile time: every
// index is a constant, so L stays in registers. A non-positive pivot makes it
// return false (the remaining columns are computed but unused).
__device__ bool cholesky_solve9(double* L, const double* rhs, double* x) {
    bool ok = true;
    double D[9], rd[9];
    sfor<0, 9>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        sfor<0, j>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const double w = L[ut(j, k)] * D[k];
            sfor<j, 9>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                L[ut(i, j)] = L[ut(i, j)] - w * L[ut(i, k)];
            });
        });
        D[j] = L[ut(j, j)];
        ok = ok && D[j] > 0.0;
        rd[j] = 1.0 / D[j];
        sfor<j + 1, 9>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            L[ut(i, j)] = L[ut(i, j)] * rd[j];
        });
    });
    if (!ok) return false;
    double y[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) y[i] = rhs[i];
    sfor<0, 9>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        sfor<i + 1, 9>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            y[k] = y[k] - L[ut(k, i)] * y[i];
        });
    });
#pragma unroll
    for (int i = 0; i < 9; ++i) y[i] = y[i] * rd[i];
    sfor<0, 9>([&](auto ic) {
        constexpr int i = 8 - decltype(ic)::value;
        double dot = 0.0;
        sfor<i + 1, 9>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            dot = dot + L[ut(k, i)] * y[k];
        });
        y[i] = y[i] - dot;
    });
#pragma unroll
    for (int i = 0; i < 9; ++i) x[i] = y[i];
    return true;
}

// ------------------------------------------------------------- model choice

__device__ __forceinline__ uint64_t ipow_adaptive(double w, int sample, double confidence,
                                                  uint64_t max_iters) {
    // matching.rs:799-811; powi as binary exponentiation
    uint64_t floor_ = 50 < max_iters ? 50 : max_iters;
    double a = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
    unsigned ub = (unsigned)sample;
    double r = 1.0;
    for (;;) {
        if (ub & 1) r *= a;
        ub /= 2;
        if (ub == 0) break;
        a *= a;
    }
    if (r >= 1.0) return floor_;
    double denom = log1p(-r);
    const double cmax = 1.0 - 1e-15;
    double cc = confidence < 0.0 ? 0.0 : (confidence > cmax ? cmax : confidence);
    double need = log1p(-cc) / denom;
    if (!isfinite(need) || need >= (double)max_iters) return max_iters;
    double ce = ceil(need);
    uint64_t v = ce <= 0.0 ? 0 : (uint64_t)ce;
    if (v < floor_) v = floor_;
    if (v > max_iters) v = max_iters;
    return v;
}

// Binomial tail P(X >= k), X ~ B(n, p) (matching.rs:843-865), for a k per
// lane; n and p are the same in every lane, and every lane of the wave must
// take part. The host loop's operands are formed 64 indices at a time, lane
// j taking index i0 + j (the ln C(n, i) increment, then the term
// exp(ln C(n, i) + i lp + (n - i) lq)); both running sums (ln C(n, i) and
// the tail) then add them in index order, as the host does, every lane
// reading them lane by lane -- so each lane's tail is bit-identical to the
// host's for its k, while the logarithms and exponentials run in parallel.
__device__ double binomial_tails(uint64_t n, uint64_t k, double p) {
    const int lane = lane_id();
    double triv = 1.0;
    bool trivial = true;
    if (k == 0)
        triv = 1.0;
    else if (k > n || p <= 0.0)
        triv = 0.0;
    else if (p >= 1.0)
        triv = 1.0;
    else
        trivial = false;
    if (__ballot(!trivial) == 0) return triv;
    const double lp = log(p), lq = log1p(-p);
    double lc = 0.0, sum = 0.0;
    for (uint64_t i0 = 0; i0 <= n; i0 += NL) {
        const uint64_t i = i0 + (uint64_t)lane;
        double inc = 0.0;
        if (i > 0 && i <= n) inc = log((double)(n - i + 1)) - log((double)i);
        const uint32_t cnt = n + 1 - i0 < (uint64_t)NL ? (uint32_t)(n + 1 - i0) : (uint32_t)NL;
        double mine = 0.0;
        for (uint32_t j = 0; j < cnt; ++j) {
            const double v = readlane_d(inc, (int)j);
            if (i0 + j > 0) lc = lc + v;
            mine = j == (uint32_t)lane ? lc : mine;
        }
        double term = 0.0;
        if (i <= n) term = exp(mine + (double)i * lp + (double)(n - i) * lq);
        for (uint32_t j = 0; j < cnt; ++j) {
            const double v = readlane_d(term, (int)j);
            if (i0 + j >= k) sum = sum + v;
        }
    }
    return trivial ? triv : (sum < 1.0 ? sum : 1.0);
}

// matching.rs:405-409: the first k in 2..=n whose tail at rate 0.001 is
// below 1e-6 (n + 1 if none), 64 candidate k at a time. Wave-uniform.
__device__ uint64_t parallax_needed(uint64_t n) {
    for (uint64_t k0 = 2; k0 <= n; k0 += NL) {
        const uint64_t k = k0 + (uint64_t)lane_id();
        const double tail = binomial_tails(n, k, 0.001);
        const uint64_t hit = __ballot(k <= n && tail < 1e-6);
        if (hit) return k0 + (uint64_t)__builtin_ctzll(hit);
    }
    return n + 1;
}

// ------------------------------------------------------ wave-cooperative sums
//
// The workgroup is one 64-lane wave. A sum over a selected point set keeps
// the host's order (ascending point index, one running sum per quantity):
// the lanes compute the per-point terms of a 64-point chunk in parallel,
// compact the selected points' terms into LDS in index order, and lane k
// then adds term k of every staged point to its own running sum. Each sum
// sees the same operands in the same order as the host loop, so every total
// is bit-identical, while the per-point work runs on all 64 lanes. Values
// that every lane computes from the same inputs are identical in all lanes.

constexpr int kStageW = 10;  // doubles staged per point

constexpr int kLmSpec = 8;  // LM tries of one iteration evaluated side by side

struct Coop {
    double stage[NL][kStageW];
    double cand[kLmSpec][9];  // the candidate F of each LM try
};
__shared__ Coop s_cp[kWaves];

__device__ __forceinline__ Coop& coop() { return s_cp[wave_id()]; }

__device__ __forceinline__ uint64_t lanes_below() {
    return (1ull << lane_id()) - 1ull;
}

// ---------------------------------------------------------------- sampler
//
// The seeded LCG streams (matching.rs:956-962, :314-320, :884-889) are
// generated 64 draws at a time: lane l jumps l + 1 steps ahead
// (st' = A^(l+1) st + c_(l+1) mod 2^64) and reduces its state to a draw;
// the rejection rule then runs on the scalar unit over the draws in order.

struct LcgLane {
    uint64_t mul, add;  // jump of lane + 1 steps
};

__device__ __forceinline__ LcgLane lcg_lane_jump() {
    uint64_t m = 1, a = 0;
    for (int j = 0; j <= lane_id(); ++j) {
        m = m * 6364136223846793005ull;
        a = a * 6364136223846793005ull + 1442695040888963407ull;
    }
    return {m, a};
}

struct LcgStream {
    uint64_t next;  // state after the last generated draw (uniform)
    uint32_t raw;   // this lane's draw of the current block of 64
    uint32_t pos;   // first unconsumed draw of the block; 64 = none left
};

__device__ __forceinline__ void lcg_refill(LcgStream& s, const LcgLane& j, uint32_t m) {
    const uint64_t st = j.mul * s.next + j.add;
    s.raw = (uint32_t)(st >> 33) % m;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)st, NL - 1);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(st >> 32), NL - 1);
    s.next = ((uint64_t)hi << 32) | lo;
    s.pos = 0;
}

// nh samples of PER distinct draws in [0, m) each, in stream order (a draw
// equal to one already taken by the sample is redrawn); lane 0 writes
// out[h * stride + e]. All lanes.
template <int PER>
__device__ void lcg_samples(LcgStream& s, const LcgLane& j, uint32_t m, int nh, uint32_t* out,
                            int stride) {
    for (int h = 0; h < nh; ++h) {
        uint32_t v[PER];
#pragma unroll
        for (int e = 0; e < PER; ++e) v[e] = 0;
        int k = 0;
        while (k < PER) {
            if (s.pos == NL) lcg_refill(s, j, m);
            const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)s.raw, (int)s.pos);
            ++s.pos;
            bool dup = false;
#pragma unroll
            for (int e = 0; e < PER; ++e) dup |= e < k && v[e] == c;
            if (!dup) {
#pragma unroll
                for (int e = 0; e < PER; ++e) v[e] = e == k ? c : v[e];
                ++k;
            }
        }
        if (lane_id() == 0) {
#pragma unroll
            for (int e = 0; e < PER; ++e) out[h * stride + e] = v[e];
        }
    }
}

// Bits [base, base + 64) of an n-bit bitset stored in 32-bit words.
__device__ __forceinline__ uint64_t bits64(const uint32_t* bits, uint32_t base, uint32_t n) {
    const uint32_t w = base >> 5;
    uint64_t m = bits[w];
    if (base + 32 < n) m |= (uint64_t)bits[w + 1] << 32;
    return m;
}

// acc = acc + stage[j][u] for the cnt staged points, in order. The loads of
// a group of eight are issued while the previous group is being added.
__device__ __forceinline__ double stage_sum(int u, uint32_t cnt, double acc) {
    uint32_t j = 0;
    if (cnt >= 8) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = coop().stage[e][u];
        for (j = 8; j + 8 <= cnt; j += 8) {
            double nx[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) nx[e] = coop().stage[j + e][u];
#pragma unroll
            for (int e = 0; e < 8; ++e) acc = acc + v[e];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = nx[e];
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) acc = acc + v[e];
    }
    for (; j < cnt; ++j) acc = acc + coop().stage[j][u];
    return acc;
}

// acc = acc + stage[j][u] * stage[j][w] for the cnt staged points, in order,
// pipelined like stage_sum.
__device__ __forceinline__ double stage_dot(int u, int w, uint32_t cnt,
                                            double acc) {
    uint32_t j = 0;
    if (cnt >= 8) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = coop().stage[e][u] * coop().stage[e][w];
        for (j = 8; j + 8 <= cnt; j += 8) {
            double a[8], b[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                a[e] = coop().stage[j + e][u];
                b[e] = coop().stage[j + e][w];
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) acc = acc + v[e];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = a[e] * b[e];
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) acc = acc + v[e];
    }
    for (; j < cnt; ++j) acc = acc + coop().stage[j][u] * coop().stage[j][w];
    return acc;
}

// The points of one index list (up to kCacheN of them) kept in the wave's LDS between
// passes: the first pass over a list gathers them from global memory and fills the cache,
// the passes that follow over the same list read it (no gathers, no L2 round trips).
constexpr uint32_t kCacheN = 128;
struct PtCache {
    double2 p[kCacheN], q[kCacheN];
};
__shared__ PtCache s_pt[kWaves];
enum CacheMode { kNoCache = 0, kFill = 1, kUse = 2 };

// One pass over the points list[0..len) in chunks of NL: stage(slot, p, q)
// for each lane holding a point of the chunk (slot = lane), then, once the
// wave has staged, reduce(cnt) over the chunk's cnt points. Without the cache,
// the next chunk's points and the indices two chunks ahead are loaded while a
// chunk is summed. kFill / kUse require len <= kCacheN.
template <class Stage, class Reduce>
__device__ __forceinline__ void list_pass(const double2* x1, const double2* x2,
                                          const uint32_t* list, uint32_t len, Stage&& stage,
                                          Reduce&& reduce, CacheMode mode = kNoCache) {
    const uint32_t lane = (uint32_t)lane_id();
    PtCache& pc = s_pt[wave_id()];
    if (mode == kUse) {
        for (uint32_t base = 0; base < len; base += NL) {
            const uint32_t cnt = len - base < (uint32_t)NL ? len - base : (uint32_t)NL;
            if (lane < cnt) stage(lane, pc.p[base + lane], pc.q[base + lane]);
            lds_sync();
            reduce(cnt);
            lds_sync();
        }
        return;
    }
    if (len == 0) return;
    // every load is issued unconditionally at an index clamped into the list
    // (lanes past the end reload its last point, unused), so no lane-masked
    // branch separates a load from the wait for it
    const uint32_t last = len - 1;
    uint32_t inext = ldg(list, lane < last ? lane : last);
    double2 p = ldg2(x1, inext), q = ldg2(x2, inext);
    inext = ldg(list, lane + NL < last ? lane + NL : last);
    for (uint32_t base = 0; base < len; base += NL) {
        const uint32_t cnt = len - base < (uint32_t)NL ? len - base : (uint32_t)NL;
        if (lane < cnt) {
            if (mode == kFill) {
                pc.p[base + lane] = p;
                pc.q[base + lane] = q;
            }
            stage(lane, p, q);
        }
        // the next chunk's points and the indices two chunks ahead load while
        // this chunk is summed
        p = ldg2(x1, inext);
        q = ldg2(x2, inext);
        const uint32_t j2 = base + 2 * NL + lane;
        inext = ldg(list, j2 < last ? j2 : last);
        lds_sync();
        reduce(cnt);
        lds_sync();
    }
}

// (row, col) of upper-triangle slot k (ut order), k < 45.
__device__ __forceinline__ void ut_rc(int k, int& r, int& c) {
    r = 0;
    while (k >= 9 - r) {
        k -= 9 - r;
        ++r;
    }
    c = r + k;
}

// Sampson errors of F over all n points, one pass: the bitset {d < th2}
// into `at` and its size (returned); when `sel` is set, the bitset {d < t2}
// into `sel` (overwritten) and its size into *ns, and when `changed` is set
// too, whether `sel` differed from the new bitset into *changed.
__device__ uint32_t coop_masks2(const double* F, const double2* x1, const double2* x2, uint32_t n,
                                double th2, uint32_t* at, double t2, uint32_t* sel,
                                uint32_t* ns, bool* changed = nullptr) {
    const int lane = lane_id();
    uint32_t cnt = 0, cs = 0;
    bool diff = false;
    // the next chunk's points are loaded while this chunk's errors are computed
    double2 p = make_double2(0.0, 0.0), q = p;
    if ((uint32_t)lane < n) {
        p = ldg2(x1, lane);
        q = ldg2(x2, lane);
    }
    for (uint32_t base = 0; base < n; base += NL) {
        const uint32_t i = base + lane, inext = i + NL;
        double2 pn = make_double2(0.0, 0.0), qn = pn;
        if (inext < n) {
            pn = ldg2(x1, inext);
            qn = ldg2(x2, inext);
        }
        bool a = false, b = false;
        if (i < n) {
            const double d = sampson_error(F, p, q);
            a = d < th2;
            b = d < t2;
        }
        p = pn;
        q = qn;
        const uint64_t ma = __ballot(a), mb = __ballot(b);
        if (lane < 2 && base + 32u * lane < n) {
            const uint32_t w = (base >> 5) + lane;
            at[w] = lane ? (uint32_t)(ma >> 32) : (uint32_t)ma;
            if (sel) {
                const uint32_t nw = lane ? (uint32_t)(mb >> 32) : (uint32_t)mb;
                if (changed) diff |= sel[w] != nw;
                sel[w] = nw;
            }
        }
        cnt += (uint32_t)__popcll(ma);
        cs += (uint32_t)__popcll(mb);
    }
    wave_sync();
    if (ns) *ns = cs;
    if (changed) *changed = __any(diff);
    return cnt;
}

__device__ __forceinline__ uint32_t coop_mask_of(const double* F, const double2* x1,
                                                 const double2* x2, uint32_t n, double t2,
                                                 uint32_t* bits) {
    return coop_masks2(F, x1, x2, n, t2, bits, 0.0, nullptr, nullptr);
}

__device__ uint32_t coop_count_of(const double* F, const double2* x1, const double2* x2,
                                  uint32_t n, double t2) {
    uint32_t cnt = 0;
    for (uint32_t base = 0; base < n; base += NL) {
        const uint32_t i = base + lane_id();
        const bool a = i < n && sampson_error(F, x1[i], x2[i]) < t2;
        cnt += (uint32_t)__popcll(__ballot(a));
    }
    return cnt;
}

// rank -> index of the points of an n-bit bitset, in index order.
// The bitset's words are loaded 64 at a time, one per lane, and handed to the
// chunks through readlane: one load latency per 2048 points.
__device__ void coop_list(const uint32_t* bits, uint32_t n, uint32_t* list) {
    const int lane = lane_id();
    const uint32_t W = words(n);
    uint32_t done = 0;
    for (uint32_t w0 = 0; w0 < W; w0 += NL) {
        const uint32_t mine = w0 + (uint32_t)lane < W ? bits[w0 + lane] : 0u;
        const uint32_t wend = W - w0 < (uint32_t)NL ? W - w0 : (uint32_t)NL;
        for (uint32_t c = 0; c < wend; c += 2) {
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)mine, (int)c);
            const uint32_t hi =
                c + 1 < wend ? (uint32_t)__builtin_amdgcn_readlane((int)mine, (int)c + 1) : 0u;
            const uint64_t m = (uint64_t)lo | ((uint64_t)hi << 32);
            if ((m >> lane) & 1u)
                list[done + (uint32_t)__popcll(m & lanes_below())] = (w0 + c) * 32u + (uint32_t)lane;
            done += (uint32_t)__popcll(m);
        }
    }
    wave_sync();
}

// collinear_and_normalizers over the bitset `sel` (len points): fills the
// Hartley normalisers t1, t2 and returns whether either image is nearly
// collinear. Leaves the points in the wave's point cache when len <= kCacheN.
__device__ bool coop_norm(const double2* x1, const double2* x2, const uint32_t* list,
                          uint32_t len, double* t1, double* t2) {
    const int lane = lane_id();
    const bool cached = len <= kCacheN;
    double acc = 0.0;
    list_pass(x1, x2, list, len,
        [&](uint32_t slot, double2 a, double2 b) {
            coop().stage[slot][0] = a.x;
            coop().stage[slot][1] = a.y;
            coop().stage[slot][2] = b.x;
            coop().stage[slot][3] = b.y;
        },
        [&](uint32_t cnt) {
            acc = stage_sum(lane & 3, cnt, acc);
        },
        cached ? kFill : kNoCache);
    const double nc = (double)(len > 0 ? len : 1);
    const double c1x = readlane_d(acc, 0) / nc, c1y = readlane_d(acc, 1) / nc;
    const double c2x = readlane_d(acc, 2) / nc, c2y = readlane_d(acc, 3) / nc;
    acc = 0.0;
    list_pass(x1, x2, list, len,
        [&](uint32_t slot, double2 a, double2 b) {
            const double ax = a.x - c1x, ay = a.y - c1y, bx = b.x - c2x, by = b.y - c2y;
            coop().stage[slot][0] = ax * ax;
            coop().stage[slot][1] = ax * ay;
            coop().stage[slot][2] = ay * ay;
            coop().stage[slot][3] = bx * bx;
            coop().stage[slot][4] = bx * by;
            coop().stage[slot][5] = by * by;
            coop().stage[slot][6] = norm2(ax, ay);
            coop().stage[slot][7] = norm2(bx, by);
        },
        [&](uint32_t cnt) {
            acc = stage_sum(lane & 7, cnt, acc);
        },
        cached ? kUse : kNoCache);
    double q[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) q[k] = readlane_d(acc, k);
    return collinear_finish(c1x, c1y, c2x, c2y, q, len, t1, t2);
}

// A^T A of the normalised 8-point system over `sel` (matching.rs:522-538);
// lane k < 45 returns upper-triangle entry k.
__device__ double coop_ata(const double2* x1, const double2* x2, const uint32_t* list,
                           uint32_t len, const double* t1, const double* t2) {
    const int lane = lane_id();
    int r, c;
    ut_rc(lane < 45 ? lane : 0, r, c);
    double acc = 0.0;
    list_pass(x1, x2, list, len,
        [&](uint32_t slot, double2 p, double2 q) {
            double pa[3] = {p.x, p.y, 1.0}, pb[3] = {q.x, q.y, 1.0}, a[3], b[3];
            mv3(t1, pa, a);
            mv3(t2, pb, b);
            double* row = coop().stage[slot];
            row[0] = b[0] * a[0];
            row[1] = b[0] * a[1];
            row[2] = b[0];
            row[3] = b[1] * a[0];
            row[4] = b[1] * a[1];
            row[5] = b[1];
            row[6] = a[0];
            row[7] = a[1];
            row[8] = 1.0;
        },
        [&](uint32_t cnt) {
            acc = stage_dot(r, c, cnt, acc);
        },
        len <= kCacheN ? kUse : kNoCache);
    return acc;
}

// ---------------------------------------------------------------- LM refine

// Sum of squared Sampson residuals of the normalised g over `sel`
// (matching.rs:620-641 summed as the host does; each squared residual is
// (b . g a)^2 / den).
__device__ double coop_lm_cost(const double* g, double s1, double s2, const double* t1,
                               const double* t2, const double2* x1, const double2* x2,
                               const uint32_t* list, uint32_t len) {
    const int lane = lane_id();
    const double q1 = s1 * s1, q2 = s2 * s2;
    double cost = 0.0;
    list_pass(x1, x2, list, len,
        [&](uint32_t slot, double2 p, double2 q) {
            double pa[3] = {p.x, p.y, 1.0}, pb[3] = {q.x, q.y, 1.0}, a[3], b[3];
            mv3(t1, pa, a);
            mv3(t2, pb, b);
            double ga[3], gtb[3];
            mv3(g, a, ga);
            mtv3(g, b, gtb);
            double den =
                q2 * (ga[0] * ga[0] + ga[1] * ga[1]) + q1 * (gtb[0] * gtb[0] + gtb[1] * gtb[1]);
            double e2 = 0.0;
            if (den > 0.0) {
                const double num = b[0] * ga[0] + b[1] * ga[1] + b[2] * ga[2];
                e2 = num * num / den;
            }
            coop().stage[slot][0] = e2;
        },
        [&](uint32_t cnt) {
            cost = stage_sum(0, cnt, cost);
        },
        len <= kCacheN ? kUse : kNoCache);
    return cost;
}

// coop_lm_cost of the kLmSpec candidates coop().cand[0..kLmSpec): lane k
// (and every lane congruent to k) returns candidate k's sum.
__device__ double coop_lm_costs(double s1, double s2, const double* t1, const double* t2,
                                const double2* x1, const double2* x2, const uint32_t* list,
                                uint32_t len) {
    const int lane = lane_id();
    const double q1 = s1 * s1, q2 = s2 * s2;
    double cost = 0.0;
    list_pass(x1, x2, list, len,
        [&](uint32_t slot, double2 p, double2 q) {
            double pa[3] = {p.x, p.y, 1.0}, pb[3] = {q.x, q.y, 1.0}, a[3], b[3];
            mv3(t1, pa, a);
            mv3(t2, pb, b);
#pragma unroll
            for (int k = 0; k < kLmSpec; ++k) {
                double g[9];
                load9(coop().cand[k], g);
                double ga[3], gtb[3];
                mv3(g, a, ga);
                mtv3(g, b, gtb);
                double den = q2 * (ga[0] * ga[0] + ga[1] * ga[1]) +
                             q1 * (gtb[0] * gtb[0] + gtb[1] * gtb[1]);
                double e2 = 0.0;
                if (den > 0.0) {
                    const double num = b[0] * ga[0] + b[1] * ga[1] + b[2] * ga[2];
                    e2 = num * num / den;
                }
                coop().stage[slot][k] = e2;
            }
        },
        [&](uint32_t cnt) {
            cost = stage_sum(lane & (kLmSpec - 1), cnt, cost);
        },
        len <= kCacheN ? kUse : kNoCache);
    return cost;
}

// J^T J (lanes 0..44, upper-triangle slots) and J^T r (lanes 45..53) of the
// Sampson residuals over `sel` (matching.rs:585-618, :690-706).
__device__ double coop_normal_eq(const double* g, double s1, double s2,
                                 const double* t1, const double* t2, const double2* x1,
                                 const double2* x2, const uint32_t* list, uint32_t len) {
    const int lane = lane_id();
    const double q1 = s1 * s1, q2 = s2 * s2;
    int u = 0, w = 0;
    if (lane < 45)
        ut_rc(lane, u, w);
    else if (lane < 54) {
        u = lane - 45;
        w = 9;
    }
    double acc = 0.0;
    list_pass(x1, x2, list, len,
        [&](uint32_t slot, double2 p, double2 q) {
            double pa[3] = {p.x, p.y, 1.0}, pb[3] = {q.x, q.y, 1.0}, a[3], b[3];
            mv3(t1, pa, a);
            mv3(t2, pb, b);
            double ga[3], gtb[3];
            mv3(g, a, ga);
            mtv3(g, b, gtb);
            double e = b[0] * ga[0] + b[1] * ga[1] + b[2] * ga[2];
            double den =
                q2 * (ga[0] * ga[0] + ga[1] * ga[1]) + q1 * (gtb[0] * gtb[0] + gtb[1] * gtb[1]);
            double* row = coop().stage[slot];
            if (den <= 0.0) {
#pragma unroll
                for (int k = 0; k < 10; ++k) row[k] = 0.0;
            } else {
                double inv = 1.0 / sqrt(den);
                double c = e * inv * inv * inv;
                double uu[3] = {q2 * ga[0], q2 * ga[1], 0.0};
                double vv[3] = {q1 * gtb[0], q1 * gtb[1], 0.0};
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int l = 0; l < 3; ++l)
                        row[3 * k + l] = b[k] * a[l] * inv - c * (uu[k] * a[l] + b[k] * vv[l]);
                row[9] = e * inv;
            }
        },
        [&](uint32_t cnt) {
            acc = stage_dot(u, w, cnt, acc);
        },
        len <= kCacheN ? kUse : kNoCache);
    return acc;
}

// refine_sampson (matching.rs:661-760) over the bitset `sel` (len points),
// all lanes; out receives the refined F (or F unchanged). list: scratch for
// the selected indices (len words).
__device__ void coop_refine_sampson(const double* F, const double2* x1,
                                                 const double2* x2, uint32_t n,
                                                 const uint32_t* sel, uint32_t len, int iters,
                                                 double* out, uint32_t* list) {
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = F[i];
    if (len < 8) return;
    coop_list(sel, n, list);
    double t1[9], t2[9], t1i[9], t2i[9];
    coop_norm(x1, x2, list, len, t1, t2);
    if (!inv3(t1, t1i) || !inv3(t2, t2i)) return;
    const double s1 = t1[0], s2 = t2[0];
    double t2it[9], tmp[9], g0[9], g[9];
    tr3(t2i, t2it);
    mm3(t2it, F, tmp);
    mm3(tmp, t1i, g0);
    // V: the right singular basis of g; each damped step's rank-2 projection starts its
    // Jacobi rotations from it (g + d is close to g)
    double V[9];
    if (!rank2_unit_v(g0, nullptr, g, V)) return;
    double cost = coop_lm_cost(g, s1, s2, t1, t2, x1, x2, list, len);
    double lambda = 1e-3;
    const int lane = lane_id();
    for (int it = 0; it < iters; ++it) {
        const double ne = coop_normal_eq(g, s1, s2, t1, t2, x1, x2, list, len);
        double jtj[45], jtr[9];
#pragma unroll
        for (int k = 0; k < 45; ++k) jtj[k] = readlane_d(ne, k);
#pragma unroll
        for (int k = 0; k < 9; ++k) jtr[k] = readlane_d(ne, 45 + k);
        // The host's up to kLmSpec tries run one after another, each failure (no
        // Cholesky factor, no rank-2 unit F, no lower cost) multiplying lambda by
        // 10. Here lane k takes try k with lambda 10^k-fold (multiplied up in the
        // same order), the costs of all tries are summed side by side, and the
        // first try in order that succeeds is the one the host accepts.
        const int k = lane & (kLmSpec - 1);
        double lam = lambda;
        for (int j = 0; j < k; ++j) lam *= 10.0;
        double m[45];
#pragma unroll
        for (int i = 0; i < 45; ++i) m[i] = jtj[i];
#pragma unroll
        for (int i = 0; i < 9; ++i) m[ut(i, i)] = m[ut(i, i)] + lam * (jtj[ut(i, i)] + 1e-12);
        double nrhs[9], d[9], cand[9], Vc[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) nrhs[i] = -jtr[i];
        bool ok = cholesky_solve9(m, nrhs, d);
        if (ok) {
            double gs[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) gs[i] = g[i] + d[i];
            ok = rank2_unit_v(gs, V, cand, Vc);
        }
        if (!ok) {
#pragma unroll
            for (int i = 0; i < 9; ++i) cand[i] = 0.0;
        }
        if (lane < kLmSpec) {
#pragma unroll
            for (int i = 0; i < 9; ++i) coop().cand[lane][i] = cand[i];
        }
        wave_sync();
        const double c = coop_lm_costs(s1, s2, t1, t2, x1, x2, list, len);
        const uint64_t win = __ballot(lane < kLmSpec && ok && c < cost);
        bool improved = false;
        if (win) {
            const int kb = __ffsll((unsigned long long)win) - 1;
            const double cb = readlane_d(c, kb), lb = readlane_d(lam, kb);
            load9(coop().cand[kb], g);
#pragma unroll
            for (int i = 0; i < 9; ++i) V[i] = readlane_d(Vc[i], kb);
            const double rel = (cost - cb) / (cost > 1e-300 ? cost : 1e-300);
            cost = cb;
            lambda = lb * 0.1 > 1e-9 ? lb * 0.1 : 1e-9;
            improved = rel > 1e-10;
        }
        wave_sync();
        if (!improved) break;
    }
    double t2t[9], fp[9];
    tr3(t2, t2t);
    mm3(t2t, g, tmp);
    mm3(tmp, t1, fp);
    double nn = fro3(fp);
    if (isfinite(nn) && nn > 0.0) {
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = fp[i] / nn;
    }
}

// ------------------------------------------------------------ homography
//
// two_view.rs:605-642 on Hartley-normalised points; matching.rs:251-369.

__device__ __forceinline__ bool htr(const double* m, double2 x, double2& out) {
    double p[3] = {x.x, x.y, 1.0}, v[3];
    mv3(m, p, v);
    if (fabs(v[2]) < 1e-12) return false;
    out.x = v[0] / v[2];
    out.y = v[1] / v[2];
    return true;
}

__device__ __forceinline__ double homography_error(const double* h, const double* hi, double2 p,
                                                   double2 q) {
    double2 a, b;
    if (!htr(h, p, a) || !htr(hi, q, b)) return INFINITY;
    return fmax(norm2(a.x - q.x, a.y - q.y), norm2(b.x - p.x, b.y - p.y));
}

// Four samples given as ranks among the inliers (`list`: rank -> index).
struct PtsRank4 {
    const double2* x1;
    const double2* x2;
    const uint32_t* list;
    const uint32_t* rank;
    static constexpr uint32_t n = 4;
    __device__ bool take(uint32_t) const { return true; }
    __device__ double2 p(uint32_t i) const { return x1[list[rank[i]]]; }
    __device__ double2 q(uint32_t i) const { return x2[list[rank[i]]]; }
};

// The sampled hypotheses' homography (matching.rs:334-337): homography_dlt of
// the four samples of `s` with its DLT system (8 equations for 8 degrees of
// freedom) solved in closed form, the host's homography_4pt step for step:
// Hartley normalisers, normalised points p_i, q_i, adj(M) rows p1 x p2,
// p2 x p0, p0 x p1 (M = [p0 p1 p2], likewise N), weights l = adj(M) p3,
// m = adj(N) q3, Hn = N diag(m_i / l_i) adj(M), then T2^-1 Hn T1 over its
// Frobenius norm. False if some l_i is 0 or a result is not finite. One lane.
__device__ __forceinline__ void cross3(const double* u, const double* v, double* w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}
template <class S>
__device__ bool homography_4pt(const S& s, double* H) {
    double t1[9], t2[9];
    normalizer(s, false, t1, 4);
    normalizer(s, true, t2, 4);
    double p[4][3], q[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double2 pp = s.p(i), qq = s.q(i);
        double p3[3] = {pp.x, pp.y, 1.0}, v[3];
        mv3(t1, p3, v);
        p[i][0] = v[0] / v[2];
        p[i][1] = v[1] / v[2];
        p[i][2] = 1.0;
        double q3[3] = {qq.x, qq.y, 1.0};
        mv3(t2, q3, v);
        q[i][0] = v[0] / v[2];
        q[i][1] = v[1] / v[2];
        q[i][2] = 1.0;
    }
    double am[3][3], an[3][3], w[3];
    cross3(p[1], p[2], am[0]);
    cross3(p[2], p[0], am[1]);
    cross3(p[0], p[1], am[2]);
    cross3(q[1], q[2], an[0]);
    cross3(q[2], q[0], an[1]);
    cross3(q[0], q[1], an[2]);
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double l = am[i][0] * p[3][0] + am[i][1] * p[3][1] + am[i][2] * p[3][2];
        const double m = an[i][0] * q[3][0] + an[i][1] * q[3][1] + an[i][2] * q[3][2];
        ok = ok && l != 0.0;
        w[i] = m / l;
    }
    if (!ok) return false;
    double hn[9];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double t0 = q[0][r] * w[0], tt1 = q[1][r] * w[1], tt2 = q[2][r] * w[2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            hn[3 * r + c] = t0 * am[0][c] + tt1 * am[1][c] + tt2 * am[2][c];
            ok = ok && isfinite(hn[3 * r + c]);
        }
    }
    if (!ok) return false;
    double t2i[9], tmp[9], h[9];
    if (!inv3(t2, t2i)) return false;
    mm3(t2i, hn, tmp);
    mm3(tmp, t1, h);
    const double nn = fro3(h);
    if (!(isfinite(nn) && nn > 0.0)) return false;
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = h[i] / nn;
    return true;
}

// Eigen-solve and denormalisation of a DLT normal matrix by a solver group.
__device__ bool homography_finish_group(double* ata, const double* t1, const double* t2,
                                        double* H) {
    double hn[9], ev[9];
    eigen_min_group(ata, ev, hn);
    bool fin = true;
#pragma unroll
    for (int i = 0; i < 9; ++i) fin = fin && isfinite(hn[i]);
    if (!fin) return false;
    double t2i[9], tmp[9], h[9];
    if (!inv3(t2, t2i)) return false;
    mm3(t2i, hn, tmp);
    mm3(tmp, t1, h);
    double nn = fro3(h);
    if (!(isfinite(nn) && nn > 0.0)) return false;
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = h[i] / nn;
    return true;
}

// r1 / r2 entries of a DLT point row as indices into its staged vector
// [-x, -y, -1, 0, u*x, u*y, u, w*x, w*y, w].
__constant__ int kDltR1[9] = {0, 1, 2, 3, 3, 3, 4, 5, 6};
__constant__ int kDltR2[9] = {3, 3, 3, 0, 1, 2, 7, 8, 9};

// homography_dlt over the bitset `sel` (len points), all lanes; H is the
// same in every lane. list: scratch for the selected indices (len words).
__device__ bool coop_homography_dlt(const double2* x1, const double2* x2, uint32_t n,
                                    const uint32_t* sel, uint32_t len, double* H,
                                    uint32_t* list) {
    const int lane = lane_id();
    if (len < 4) return false;
    coop_list(sel, n, list);
    double t1[9], t2[9];
    coop_norm(x1, x2, list, len, t1, t2);
    int r, c;
    ut_rc(lane < 45 ? lane : 0, r, c);
    const int a1 = kDltR1[r], b1 = kDltR1[c], a2 = kDltR2[r], b2 = kDltR2[c];
    double acc = 0.0;
    list_pass(x1, x2, list, len,
        [&](uint32_t slot, double2 pp, double2 qq) {
            double p3[3] = {pp.x, pp.y, 1.0}, v[3];
            mv3(t1, p3, v);
            const double x = v[0] / v[2], y = v[1] / v[2];
            double q3[3] = {qq.x, qq.y, 1.0};
            mv3(t2, q3, v);
            const double u = v[0] / v[2], w = v[1] / v[2];
            double* row = coop().stage[slot];
            row[0] = -x;
            row[1] = -y;
            row[2] = -1.0;
            row[3] = 0.0;
            row[4] = u * x;
            row[5] = u * y;
            row[6] = u;
            row[7] = w * x;
            row[8] = w * y;
            row[9] = w;
        },
        [&](uint32_t cnt) {
            uint32_t j = 0;
            for (; j + 8 <= cnt; j += 8) {
                double v[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const double* row = coop().stage[j + e];
                    v[e] = row[a1] * row[b1] + row[a2] * row[b2];
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) acc = acc + v[e];
            }
            for (; j < cnt; ++j) {
                const double* row = coop().stage[j];
                acc = acc + (row[a1] * row[b1] + row[a2] * row[b2]);
            }
        },
        len <= kCacheN ? kUse : kNoCache);
    double ata[45];
#pragma unroll
    for (int k = 0; k < 45; ++k) ata[k] = readlane_d(acc, k);
    // every solver group solves the same system; lane 0's result is used
    double h[9];
    const int ok = homography_finish_group(ata, t1, t2, h) ? 1 : 0;
    if (!__builtin_amdgcn_readlane(ok, 0)) return false;
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = readlane_d(h[i], 0);
    return true;
}

// |{i in sel : homography_error(h, hi, i) < th}|; with `out`, also the
// bitset of those points.
__device__ uint32_t coop_count_h(const double* h, const double* hi, const double2* x1,
                                 const double2* x2, uint32_t n, const uint32_t* sel, double th,
                                 uint32_t* out) {
    const int lane = lane_id();
    uint32_t cnt = 0;
    for (uint32_t base = 0; base < n; base += NL) {
        const uint64_t m = bits64(sel, base, n);
        const bool in = ((m >> lane) & 1u) &&
                        homography_error(h, hi, x1[base + lane], x2[base + lane]) < th;
        const uint64_t b = __ballot(in);
        if (out && lane < 2 && base + 32u * lane < n)
            out[(base >> 5) + lane] = lane ? (uint32_t)(b >> 32) : (uint32_t)b;
        cnt += (uint32_t)__popcll(b);
    }
    if (out) wave_sync();
    return cnt;
}

constexpr int kHypPerBatch = 3 * kGroups;  // hypotheses per batch: one per solver group, 3 rounds

struct SelState {
    double F[kHypPerBatch][9];
    int cnt[kHypPerBatch];
    uint32_t idx[kHypPerBatch][4];
    double best[9];
    int have, bcnt;
    uint64_t iters, it;
};

// Solver group g fits homography hypothesis kGroups * round + g from its
// four samples and counts its inliers over the ns listed points (split
// across the group's lanes).
__device__ void hom_round(const double2* x1, const double2* x2, const uint32_t* list,
                          uint32_t ns, double th, SelState& sh, int round) {
    const int g = solver_group(), part = solver_part(), gbase = 3 * g;
    const int h = kGroups * round + g;
    if (g >= kGroups) return;
    PtsRank4 s4{x1, x2, list, sh.idx[h]};
    double hh[9], hi[9];
    int c = -1;
    if (homography_4pt(s4, hh) && inv3(hh, hi)) {
        // ranks part, part + 3, ...: four per step, loads issued together
        uint32_t cc = 0, r = part;
        for (; r + 9 < ns; r += 12) {
            const uint32_t i0 = list[r], i1 = list[r + 3], i2 = list[r + 6], i3 = list[r + 9];
            const double2 a0 = x1[i0], b0 = x2[i0], a1 = x1[i1], b1 = x2[i1];
            const double2 a2 = x1[i2], b2 = x2[i2], a3 = x1[i3], b3 = x2[i3];
            cc += homography_error(hh, hi, a0, b0) < th ? 1u : 0u;
            cc += homography_error(hh, hi, a1, b1) < th ? 1u : 0u;
            cc += homography_error(hh, hi, a2, b2) < th ? 1u : 0u;
            cc += homography_error(hh, hi, a3, b3) < th ? 1u : 0u;
        }
        for (; r < ns; r += 3) {
            const uint32_t i = list[r];
            cc += homography_error(hh, hi, x1[i], x2[i]) < th ? 1u : 0u;
        }
        c = (int)(__shfl(cc, gbase) + __shfl(cc, gbase + 1) + __shfl(cc, gbase + 2));
        if (part == 0) {
#pragma unroll
            for (int i = 0; i < 9; ++i) sh.F[h][i] = hh[i];
        }
    }
    if (part == 0) sh.cnt[h] = c;
}

__constant__ double kHM[4] = {4.0, 2.0, 1.0, 1.0};

// matching.rs:300-369 over the inlier set `sel` (ns points, `list` maps
// their ranks to indices), all lanes. H, Hi, *cnt_out are the same in every
// lane; selbits (words(n) words) and list_buf (n words) are scratch.
__device__ bool fit_homography(const double2* x1, const double2* x2, uint32_t n,
                                            const uint32_t* sel, const uint32_t* list,
                                            uint32_t ns, double th, double* H, double* Hi,
                                            uint32_t* cnt_out, SelState& sh, uint32_t* selbits,
                                            uint32_t* list_buf) {
    const int lane = lane_id();
    if (ns < 4) return false;
    if (lane == 0) {
        sh.have = 0;
        sh.bcnt = 0;
        sh.iters = 500;
        sh.it = 0;
    }
    wave_sync();
    const LcgLane jump = lcg_lane_jump();
    LcgStream hs{0x2545F4914F6CDD1Dull, 0u, (uint32_t)NL};
    while (sh.it < sh.iters) {
        // matching.rs:314-320: four distinct ranks among the inliers
        lcg_samples<4>(hs, jump, ns, kHypPerBatch, &sh.idx[0][0], 4);
        wave_sync();
        for (int round = 0; round < 3; ++round) hom_round(x1, x2, list, ns, th, sh, round);
        wave_sync();
        if (lane == 0) {
            for (int h = 0; h < kHypPerBatch && sh.it < sh.iters; ++h) {
                sh.it++;
                const int c = sh.cnt[h];
                if (c < 0) continue;
                if (!sh.have || c > sh.bcnt) {
                    sh.have = 1;
                    sh.bcnt = c;
#pragma unroll
                    for (int i = 0; i < 9; ++i) sh.best[i] = sh.F[h][i];
                    sh.iters = ipow_adaptive((double)c / (double)ns, 4, 0.999, 500);
                }
            }
        }
        wave_sync();
    }
    if (!sh.have) return false;
    double bh[9], bhi[9], ch[9], chi[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) bh[i] = sh.best[i];
    inv3(bh, bhi);
    uint32_t bc = (uint32_t)sh.bcnt;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        ch[i] = bh[i];
        chi[i] = bhi[i];
    }
    for (int r = 0; r < 4; ++r) {
        const uint32_t nsel = coop_count_h(ch, chi, x1, x2, n, sel, th * kHM[r], selbits);
        double gg[9], gi[9];
        if (!coop_homography_dlt(x1, x2, n, selbits, nsel, gg, list_buf)) break;
        if (!inv3(gg, gi)) break;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            ch[i] = gg[i];
            chi[i] = gi[i];
        }
        const uint32_t c = coop_count_h(gg, gi, x1, x2, n, sel, th, nullptr);
        if (c > bc) {
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                bh[i] = gg[i];
                bhi[i] = gi[i];
            }
            bc = c;
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        H[i] = bh[i];
        Hi[i] = bhi[i];
    }
    *cnt_out = bc;
    return true;
}

// Hypothesis-batch state; its LDS is reused by the homography fit and the
// line test after the hypothesis loop.
struct HypMem {
    int hCnt[NL];
    double loF[NL][9];
    int loCnt[NL];
    uint32_t loNs[NL];
};
union PhaseMem {
    HypMem hyp;
    SelState hom;
    uint32_t draws[600];
};
__shared__ PhaseMem s_pm[kWaves];

// Occupancy: two waves per SIMD hide the FP64 dependency latency of the
// serial solver chains; correspondences and bitsets stay in L2-resident
// device memory and every function fits 256 VGPRs.
#ifndef SKH_RANSAC_WAVES_PER_EU
#define SKH_RANSAC_WAVES_PER_EU 2
#endif
// bitsets of W words: lo, alt, sel per hypothesis, pool (6), final (2), then
// one index list of up to 32 W >= n words per wave
constexpr uint32_t kMaskSets = 3 * NL + 8 + 32 * kWaves;

struct Shared {
    double poolF[6][9];
    int poolCnt[6];
    int poolSlot[6];
    int npool;
    double bestF[9];
    int bestCnt, haveBest;
    uint64_t iters, it;
    double stF[5][6][9];
    int stCnt[5][6];
    double f[9];
    int inlCnt;
    int result;
    uint64_t cmask;
    int pre_drawn;
    uint32_t pre_next;
    int lineBest[NL];
    double es_med;
    int es_n;
    double redA[NL];
};

// Hypotheses of one batch in device scratch: samples, F and raw counts.
struct PreBatch {
    uint32_t idx[kHypPerBatch][8];
    int cnt[64];
    double F[kHypPerBatch][9];
};
constexpr uint32_t kPreWords = (uint32_t)((sizeof(PreBatch) + 255) / 256) * 64;  // 256-byte multiple

// One round of the solver groups (lanes 3g..3g+2 form group g): groups
// [0, nlo) take the LO candidate records hand[g] (A^T A in slots 0..44, the
// Hartley scales and offsets in 45..50; results: slot 51 = 1 / 0 for Some /
// None, 52..60 = F), the next hcount groups hypotheses hfirst, hfirst + 1,
// ... of `pre` (8-point system from the samples, then the Sampson count split
// over the group's lanes; results: pre->cnt[h], -1 for None, and
// pre->F[h]). Both kinds share one solve, so a mixed round costs one solve.
// Inlined like every routine of the kernel: as out-of-line calls their frames
// cost 247 spilled VGPRs around the call sites (2032 bytes of scratch per
// lane against 1376 inlined; RANSAC stage 45.0 -> 43.0 ms on the 4000-pair
// uneven batch).
__device__ void group_solve(const double2* x1, const double2* x2, uint32_t n, double th2,
                            double* hand, int nlo, PreBatch* pre, int hfirst, int hcount) {
    const int g = solver_group(), part = solver_part(), gbase = 3 * g;
    if (g >= kGroups) return;
    const bool is_lo = g < nlo;
    if (!is_lo && g - nlo >= hcount) return;
    const int h = hfirst + g - nlo;
    double* hd = hand + (size_t)g * 64;
    double ata[45], t1[9], t2[9], F[9];
    bool ok = true;
    if (is_lo) {
#pragma unroll
        for (int k = 0; k < 45; ++k) ata[k] = hd[k];
        t1[0] = hd[45]; t1[1] = 0.0; t1[2] = hd[46];
        t1[3] = 0.0; t1[4] = hd[45]; t1[5] = hd[47];
        t1[6] = 0.0; t1[7] = 0.0; t1[8] = 1.0;
        t2[0] = hd[48]; t2[1] = 0.0; t2[2] = hd[49];
        t2[3] = 0.0; t2[4] = hd[48]; t2[5] = hd[50];
        t2[6] = 0.0; t2[7] = 0.0; t2[8] = 1.0;
    } else {
        PtsIdx8 s{x1, x2, pre->idx[h]};
        ok = accumulate_8pt(s, 8, t1, t2, ata);
    }
    ok = ok && solve_8pt_group(ata, t1, t2, F);
    if (is_lo) {
        if (part == 0) {
            hd[51] = ok ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < 9; ++k) hd[52 + k] = F[k];
        }
        return;
    }
    int cnt = -1;
    if (ok) {
        // points part, part + 3, ...: four per step, loads issued together
        uint32_t c = 0, i = part;
        for (; i + 9 < n; i += 12) {
            const double2 a0 = x1[i], b0 = x2[i], a1 = x1[i + 3], b1 = x2[i + 3];
            const double2 a2 = x1[i + 6], b2 = x2[i + 6], a3 = x1[i + 9], b3 = x2[i + 9];
            c += sampson_error(F, a0, b0) < th2 ? 1u : 0u;
            c += sampson_error(F, a1, b1) < th2 ? 1u : 0u;
            c += sampson_error(F, a2, b2) < th2 ? 1u : 0u;
            c += sampson_error(F, a3, b3) < th2 ? 1u : 0u;
        }
        for (; i < n; i += 3) c += sampson_error(F, x1[i], x2[i]) < th2 ? 1u : 0u;
        cnt = (int)(__shfl(c, gbase) + __shfl(c, gbase + 1) + __shfl(c, gbase + 2));
        if (part == 0) {
#pragma unroll
            for (int k = 0; k < 9; ++k) pre->F[h][k] = F[k];
        }
    }
    if (part == 0) pre->cnt[h] = cnt;
}

constexpr uint32_t kHandStride = 64;                         // doubles per candidate record
constexpr uint32_t kHandWords = 2 * kGroups * kHandStride;   // 32-bit words

// Local optimisation (matching.rs:993-1015) of every candidate in `cand`
// (hypothesis lanes), stage by stage: for each stage the candidates' point
// sums run wave-cooperatively one candidate after another, then the solver
// groups (three lanes each) solve up to 21 candidates' 8-point systems at once, then each
// candidate's Sampson pass gives its count at x1 and its next selection.
// Results: loF[h], loCnt[h] and the inlier bitset lo + h * W; list (n
// words) and hand (kGroups candidate records) are device scratch. Solver
// groups that a round leaves free solve the next batch's hypotheses of `pre`
// (reserved through the counter *pre_next).
__device__ void local_opt_batch(const RansacTask& t, HypMem& hm, uint64_t cand,
                                            double th2, uint32_t* lo, uint32_t* alt,
                                            uint32_t* selb, uint32_t W, uint32_t* list,
                                            double* hand, PreBatch* pre, uint32_t* pre_next) {
    const double2* x1 = t.x1;
    const double2* x2 = t.x2;
    const uint32_t n = t.n;
    const int lane = lane_id();
    // stage 0 selection (x3) and the x1 mask of the hypothesis itself
    for (uint64_t m = cand; m; m &= m - 1) {
        const int h = __ffsll((unsigned long long)m) - 1;
        double F0[9];
        load9(hm.loF[h], F0);
        uint32_t ns = 0;
        coop_masks2(F0, x1, x2, n, th2, lo + (size_t)h * W, th2 * kNM[0] * kNM[0],
                    selb + (size_t)h * W, &ns);
        if (lane == 0) {
            hm.loCnt[h] = hm.hCnt[h];
            hm.loNs[h] = ns;
        }
    }
    wave_sync();
    uint64_t active = cand, swapped = 0;
    for (int r = 0; r < 6 && active; ++r) {
        const double t2n = r < 5 ? th2 * kNM[r + 1] * kNM[r + 1] : 0.0;
        uint64_t todo = active;
        while (todo) {
            // up to kGroups candidates, one solver group each
            uint64_t filled = 0;
            int rank = 0;
            while (todo && rank < kGroups) {
                const int h = __ffsll((unsigned long long)todo) - 1;
                todo &= todo - 1;
                const uint32_t ns = hm.loNs[h];
                double t1[9], t2[9];
                // fundamental_8pt: fewer than 8 points or collinear -> None, LO stops
                if (ns < 8) {
                    active &= ~(1ull << h);
                    continue;
                }
                coop_list(selb + (size_t)h * W, n, list);
                if (coop_norm(x1, x2, list, ns, t1, t2)) {
                    active &= ~(1ull << h);
                    continue;
                }
                const double a = coop_ata(x1, x2, list, ns, t1, t2);
                double* hd = hand + (size_t)rank * kHandStride;
                if (lane < 45) hd[lane] = a;
                if (lane == 0) {
                    hd[45] = t1[0];
                    hd[46] = t1[2];
                    hd[47] = t1[5];
                    hd[48] = t2[0];
                    hd[49] = t2[2];
                    hd[50] = t2[5];
                }
                filled |= 1ull << h;
                ++rank;
            }
            // the groups left free take the next batch's hypotheses, in order
            uint32_t hfirst = 0;
            if (lane == 0) hfirst = atomicAdd(pre_next, (uint32_t)(kGroups - rank));
            hfirst = (uint32_t)__shfl((int)hfirst, 0);
            const int hcount = hfirst < (uint32_t)kHypPerBatch
                                   ? min(kGroups - rank, kHypPerBatch - (int)hfirst)
                                   : 0;
            wave_sync();
            group_solve(x1, x2, n, th2, hand, rank, pre, (int)hfirst, hcount);
            wave_sync();
            rank = 0;
            for (uint64_t m = filled; m; m &= m - 1, ++rank) {
                const int h = __ffsll((unsigned long long)m) - 1;
                const double* hd = hand + (size_t)rank * kHandStride;
                if (hd[51] == 0.0) {
                    active &= ~(1ull << h);
                    continue;
                }
                double gh[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) gh[k] = hd[52 + k];
                const bool sw = (swapped >> h) & 1u;
                uint32_t* spare = (sw ? lo : alt) + (size_t)h * W;
                uint32_t ns = 0;
                bool changed = true;
                const uint32_t gc = coop_masks2(gh, x1, x2, n, th2, spare, t2n,
                                                r < 5 ? selb + (size_t)h * W : nullptr, &ns,
                                                &changed);
                const int cur = hm.loCnt[h];
                wave_sync();
                if (lane == 0) hm.loNs[h] = ns;
                if ((int)gc > cur) {
                    if (lane == 0) {
#pragma unroll
                        for (int k = 0; k < 9; ++k) hm.loF[h][k] = gh[k];
                        hm.loCnt[h] = (int)gc;
                    }
                    swapped ^= 1ull << h;
                }
                // From LO stage 2 on the selection thresholds of this and every later stage
                // are equal (kNM[3..5] = 1): an unchanged selection makes the next stage solve
                // this stage's system again, giving the same F, a count that cannot beat the
                // current one and the same selection, so every remaining stage is a no-op.
                if (r >= 2 && !changed) active &= ~(1ull << h);
                wave_sync();
            }
        }
    }
    // the current inlier bitset of a swapped candidate sits in its alt slot
    for (uint64_t m = swapped; m; m &= m - 1) {
        const int h = __ffsll((unsigned long long)m) - 1;
        for (uint32_t w0 = lane; w0 < W; w0 += NL) lo[(size_t)h * W + w0] = alt[(size_t)h * W + w0];
    }
    wave_sync();
}

__shared__ Shared s_sh[kWaves];
__shared__ uint64_t s_pc[kWaves][11];

// ransac_fundamental of one pair. Small mode (big = false): the calling wave
// alone, with the LDS slot of its wave. Large mode (big = true): all waves of
// the workgroup on one pair, with LDS slot 0: wave 0 draws the samples,
// replays the batches and runs the final stages; the hypothesis rounds, the
// local optimisations of a batch and the pool refits are spread over the
// waves, each with its own staging LDS, index list and hand-off records.
__device__ void run_pair(RansacTask t, const RansacCfgDev& cfg, const bool big) {
    const int lane = lane_id(), wave = wave_id();
    const int slot = big ? 0 : wave;
    const bool lead = !big || wave == 0;
    Shared& sh = s_sh[slot];
    PhaseMem& pm = s_pm[slot];
    uint64_t* pc = s_pc[slot];
    auto pair_sync = [&]() {
        if (big)
            __syncthreads();
        else
            wave_sync();
    };
    if (t.n_dev) t.n = min(*t.n_dev, t.n);
    const uint32_t n = t.n;
    const uint32_t W = words(n);
    // device scratch: kWaves hand-off record sets, the batch hypotheses, the bitsets,
    // kWaves index lists
    double* hand = reinterpret_cast<double*>(t.scratch + (size_t)wave * kHandWords);
    PreBatch* pre = reinterpret_cast<PreBatch*>(t.scratch + (size_t)kWaves * kHandWords);
    uint32_t* masks = t.scratch + (size_t)kWaves * kHandWords + kPreWords;
    uint32_t* lo_masks = masks;                                 // NL * W
    uint32_t* alt_masks = masks + (size_t)NL * W;               // NL * W
    uint32_t* sel_masks = masks + (size_t)2 * NL * W;           // NL * W
    uint32_t* pool_masks = masks + (size_t)3 * NL * W;          // 6 * W
    uint32_t* fin_mask = masks + (size_t)(3 * NL + 6) * W;      // W
    uint32_t* fin_sel = fin_mask + W;                           // W
    uint32_t* list_buf = fin_sel + W + (size_t)wave * 32 * W;   // 32 * W >= n per wave

    // matching.rs:946-949
    bool valid = n >= 8;
    for (uint32_t i = lane; valid && i < n; i += NL) {
        double2 a = t.x1[i], b = t.x2[i];
        if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(b.x) || !isfinite(b.y)) valid = false;
    }
    valid = __all(valid);  // every wave scans every point: the same verdict
    if (!valid) {
        if (lead && lane == 0) {
            *t.found = 0;
            if (t.iters_out) *t.iters_out = 0;  // no hypothesis drawn
        }
        return;
    }
    const double th2 = cfg.threshold_px * cfg.threshold_px;
    const LcgLane jump = lcg_lane_jump();
    LcgStream smp{cfg.seed ^ 0x9E3779B97F4A7C15ull, 0u, (uint32_t)NL};
    if (lead && lane == 0) {
        sh.npool = 0;
        sh.bestCnt = 0;
        sh.haveBest = 0;
        sh.iters = cfg.max_iters;
        sh.it = 0;
        sh.result = 0;
        sh.pre_drawn = 0;
        sh.pre_next = 0;
        for (int k = 0; k < 6; ++k) sh.poolSlot[k] = k;
    }
    if (lead && lane < 11) pc[lane] = 0;
    pair_sync();

    // Phase clock (lane 0, constant-rate wall clock): sampling, hypotheses,
    // local optimisation, replay, pool LM, final refits + gates, line test,
    // model selection, batches, LO candidates, total.
    const uint64_t t_start = wall_clock64();
    uint64_t t_prev = t_start;
    auto lap = [&](int k) {
        const uint64_t now = wall_clock64();
        if (lead && lane == 0) pc[k] += now - t_prev;
        t_prev = now;
    };
    auto flush_prof = [&]() {
        if (lead && lane == 0 && t.prof) {
            pc[10] = wall_clock64() - t_start;
            for (int k = 0; k < 11; ++k) t.prof[k] = pc[k];
        }
    };

    // ---------------------------------------------------- hypothesis batches
    while (sh.it < sh.iters) {
        lap(3);
        if (lead) {
            if (lane == 0) pc[8] += 1;
            // matching.rs:971-979: eight distinct draws per hypothesis (unless the previous
            // batch already drew them)
            if (!sh.pre_drawn) {
                lcg_samples<8>(smp, jump, n, kHypPerBatch, &pre->idx[0][0], 8);
                if (lane == 0) sh.pre_next = 0;
            }
        }
        pair_sync();
        lap(0);
        {
            // the batch's hypotheses that the previous batch's LO rounds did not solve
            const int done = min((int)sh.pre_next, kHypPerBatch);
            const int rounds = (kHypPerBatch - done + kGroups - 1) / kGroups;
            for (int r = big ? wave : 0; r < rounds; r += big ? kWaves : 1) {
                const int first = done + kGroups * r;
                group_solve(t.x1, t.x2, n, th2, hand, 0, pre, first,
                            min(kGroups, kHypPerBatch - first));
            }
        }
        pair_sync();
        lap(1);
        if (lead) {
            const int cnt = lane < kHypPerBatch ? pre->cnt[lane] : -1;
            pm.hyp.hCnt[lane] = cnt;
            if (cnt >= 0) {
#pragma unroll
                for (int k = 0; k < 9; ++k) pm.hyp.loF[lane][k] = pre->F[lane][k];
            }
            // LO candidates: bound = max(best count now, raw counts >= 8 earlier in the batch)
            int v = cnt >= 8 ? cnt : 0;
            int incl = v;
#pragma unroll
            for (int off = 1; off < NL; off <<= 1) {
                int o = __shfl_up(incl, off);
                if (lane >= off) incl = incl > o ? incl : o;
            }
            int excl = __shfl_up(incl, 1);
            if (lane == 0) excl = 0;
            const int bound = sh.bestCnt > excl ? sh.bestCnt : excl;
            const bool cand = cnt >= 8 && 4 * cnt >= bound;
            const uint64_t cm = __ballot(cand);
            if (lane == 0) {
                pc[9] += __popcll(cm);
                sh.cmask = cm;
            }
            if (!cand) pm.hyp.loCnt[lane] = -1;
            // the next batch's samples: the stream continues where this batch's ended; free
            // solver groups of this batch's LO rounds solve them
            wave_sync();
            lcg_samples<8>(smp, jump, n, kHypPerBatch, &pre->idx[0][0], 8);
            if (lane == 0) {
                sh.pre_next = 0;
                sh.pre_drawn = 1;
            }
        }
        pair_sync();
        // a large pair's waves take every kWaves-th candidate
        uint64_t mine = sh.cmask;
        if (big) {
            uint64_t sub = 0;
            int r = 0;
            for (uint64_t m = mine; m; m &= m - 1, ++r)
                if (r % kWaves == wave) sub |= m & (~m + 1);
            mine = sub;
        }
        local_opt_batch(t, pm.hyp, mine, th2, lo_masks, alt_masks, sel_masks, W, list_buf, hand,
                        pre, &sh.pre_next);
        pair_sync();
        lap(2);
        // replay in order: the lead wave walks the batch's hypotheses with
        // wave-uniform control flow (the loop state in registers, every lane
        // holding the same values); the Jaccard popcounts run over the bitset
        // words one per lane, lane 0 alone updates the pool order in LDS.
        if (lead) {
            uint64_t it = sh.it, iters = sh.iters;
            int best_cnt = sh.haveBest ? sh.bestCnt : 0;
            int have_best = sh.haveBest, npool = sh.npool, result = sh.result;
            for (int h = 0; h < kHypPerBatch && it < iters; ++h) {
                it++;
                const int c0 = pm.hyp.hCnt[h];
                if (c0 < 0) continue;
                if (!(c0 >= 8 && 4 * c0 >= best_cnt)) continue;
                // loCnt[h] >= 0 here: the candidate bound never exceeds best_cnt
                const int c = pm.hyp.loCnt[h];
                if (c < 0) {  // cannot happen: candidates are a superset of triggers
                    result = -1;
                    it = iters;
                    break;
                }
                const uint32_t* inl = lo_masks + (size_t)h * W;
                int same = -1;
                for (int j = 0; j < npool && same < 0; ++j) {
                    const uint32_t* pmk = pool_masks + (size_t)sh.poolSlot[j] * W;
                    uint32_t both = 0, either = 0;
                    for (uint32_t w0 = (uint32_t)lane; w0 < W; w0 += NL) {
                        both += (uint32_t)__popc(pmk[w0] & inl[w0]);
                        either += (uint32_t)__popc(pmk[w0] | inl[w0]);
                    }
#pragma unroll
                    for (int off = NL / 2; off > 0; off >>= 1) {
                        both += (uint32_t)__shfl_xor((int)both, off);
                        either += (uint32_t)__shfl_xor((int)either, off);
                    }
                    if (10 * both > 7 * either) same = j;
                }
                int target = -1;
                if (same >= 0) {
                    if (sh.poolCnt[same] < c) target = same;
                } else {
                    target = npool++;
                }
                if (target >= 0) {
                    if (lane < 9) sh.poolF[target][lane] = pm.hyp.loF[h][lane];
                    if (lane == 0) sh.poolCnt[target] = c;
                    uint32_t* dst = pool_masks + (size_t)sh.poolSlot[target] * W;
                    for (uint32_t w0 = (uint32_t)lane; w0 < W; w0 += NL) dst[w0] = inl[w0];
                }
                wave_sync();
                if (lane == 0) {
                    // stable insertion sort, count descending
                    for (int a = 1; a < npool; ++a) {
                        double xf[9];
#pragma unroll
                        for (int i = 0; i < 9; ++i) xf[i] = sh.poolF[a][i];
                        const int xc = sh.poolCnt[a], xs = sh.poolSlot[a];
                        int b = a - 1;
                        while (b >= 0 && sh.poolCnt[b] < xc) {
#pragma unroll
                            for (int i = 0; i < 9; ++i) sh.poolF[b + 1][i] = sh.poolF[b][i];
                            sh.poolCnt[b + 1] = sh.poolCnt[b];
                            sh.poolSlot[b + 1] = sh.poolSlot[b];
                            --b;
                        }
#pragma unroll
                        for (int i = 0; i < 9; ++i) sh.poolF[b + 1][i] = xf[i];
                        sh.poolCnt[b + 1] = xc;
                        sh.poolSlot[b + 1] = xs;
                    }
                }
                wave_sync();
                if (npool > 5) npool = 5;  // the dropped slot stays at index 5
                if (c <= best_cnt) continue;
                iters = ipow_adaptive((double)c / (double)n, 8, cfg.confidence, cfg.max_iters);
                if (lane < 9) sh.bestF[lane] = pm.hyp.loF[h][lane];
                best_cnt = c;
                have_best = 1;
            }
            wave_sync();
            if (lane == 0) {
                sh.it = it;
                sh.iters = iters;
                sh.bestCnt = best_cnt;
                sh.haveBest = have_best;
                sh.npool = npool;
                sh.result = result;
            }
        }
        pair_sync();
    }
    lap(3);
    if (lead && lane == 0 && t.iters_out) *t.iters_out = sh.it;
    if (sh.result < 0) {
        if (lead && lane == 0) *t.found = -1;
        flush_prof();
        return;
    }
    if (!sh.haveBest) {
        if (lead && lane == 0) *t.found = 0;
        flush_prof();
        return;
    }

    // ------------------------------------------- pool LM (matching.rs:1041-1057)
    {
        // a large pair's waves take every kWaves-th pool entry
        const int npool = sh.npool;
        for (int p = big ? wave : 0; p < npool; p += big ? kWaves : 1) {
            uint32_t* sel = sel_masks + (size_t)p * W;
            uint32_t* cnt_bits = alt_masks + (size_t)p * W;
            double cur[9];
            load9(sh.poolF[p], cur);
            uint32_t ns = coop_mask_of(cur, t.x1, t.x2, n, th2 * kNM[0] * kNM[0], sel);
            for (int r = 0; r < 6; ++r) {
                double nxt[9];
                coop_refine_sampson(cur, t.x1, t.x2, n, sel, ns, 30, nxt, list_buf);
                load9(nxt, cur);
                if (lane == 0) {
#pragma unroll
                    for (int k = 0; k < 9; ++k) sh.stF[p][r][k] = nxt[k];
                }
                // this stage's inlier count and the next stage's selection, one pass
                const double t2n = r < 5 ? th2 * kNM[r + 1] * kNM[r + 1] : 0.0;
                const uint32_t c = coop_masks2(cur, t.x1, t.x2, n, th2, cnt_bits, t2n,
                                               r < 5 ? sel : nullptr, &ns);
                if (lane == 0) sh.stCnt[p][r] = (int)c;
            }
        }
    }
    pair_sync();
    if (!lead) return;  // the remaining stages run on the lead wave alone
    lap(4);
    {
        double f[9];
        load9(sh.bestF, f);
        int ic = sh.bestCnt;
        for (int p = 0; p < sh.npool; ++p)
            for (int r = 0; r < 6; ++r)
                if (sh.stCnt[p][r] > ic) {
                    load9(sh.stF[p][r], f);
                    ic = sh.stCnt[p][r];
                }
        // matching.rs:1059-1069
        for (int r = 0; r < 2; ++r) {
            const uint32_t ns = coop_mask_of(f, t.x1, t.x2, n, th2, fin_sel);
            double g[9];
            coop_refine_sampson(f, t.x1, t.x2, n, fin_sel, ns, 30, g, list_buf);
            const uint32_t c = coop_count_of(g, t.x1, t.x2, n, th2);
            if (c < ns) break;
            load9(g, f);
            ic = (int)c;
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 9; ++k) sh.f[k] = f[k];
        }
        coop_mask_of(f, t.x1, t.x2, n, th2, fin_mask);
        // matching.rs:1071-1080
        bool ok = true;
        const uint64_t minc = cfg.min_inliers > 8 ? cfg.min_inliers : 8;
        if ((uint64_t)ic < minc || (double)ic < cfg.min_inlier_ratio * (double)n) ok = false;
        if (ok && cfg.max_chance_prob > 0.0) {
            double lox = INFINITY, loy = INFINITY, hix = -INFINITY, hiy = -INFINITY;
            for (uint32_t i = lane; i < n; i += NL) {
                const double2 q = t.x2[i];
                lox = fmin(lox, q.x);
                loy = fmin(loy, q.y);
                hix = fmax(hix, q.x);
                hiy = fmax(hiy, q.y);
            }
            // min / max of finite values: any order gives the host's result
#pragma unroll
            for (int off = 1; off < NL; off <<= 1) {
                lox = fmin(lox, __shfl_xor(lox, off));
                loy = fmin(loy, __shfl_xor(loy, off));
                hix = fmax(hix, __shfl_xor(hix, off));
                hiy = fmax(hiy, __shfl_xor(hiy, off));
            }
            double dx = hix - lox, dy = hiy - loy;
            double area = dx * dy;
            double p = 1.0;
            if (isfinite(area) && area > 0.0) {
                p = 2.0 * cfg.threshold_px * norm2(dx, dy) / area;
                p = p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);
            }
            double chance = (uint64_t)ic <= 8 ? 1.0
                                              : binomial_tails(n > 8 ? n - 8 : 0, (uint64_t)ic - 8, p);
            if (chance > cfg.max_chance_prob) ok = false;
        }
        if (ok) {
            double t1[9], t2[9];
            coop_list(fin_mask, n, list_buf);
            if (coop_norm(t.x1, t.x2, list_buf, (uint32_t)ic, t1, t2)) ok = false;
        }
        if (lane == 0) {
            sh.result = ok ? 1 : 0;
            sh.es_n = ic;
        }
    }
    wave_sync();
    lap(5);
    if (!sh.result) {
        if (lane == 0) *t.found = 0;
        flush_prof();
        return;
    }

    // ---------------------------- off_line_count (matching.rs:878-906), both images
    const uint32_t ns = (uint32_t)sh.es_n;
    // rank -> index of the final inliers (the samplers below draw ranks)
    uint32_t* inl_list = alt_masks;
    coop_list(fin_mask, n, inl_list);
    bool line_ok = true;
    for (int side = 0; side < 2; ++side) {
        if (ns < 3) {
            line_ok = false;  // off_line_count = 0 < 8
            break;
        }
        {
            // matching.rs:884-889: 300 pairs of draws
            LcgStream ls{0x6A09E667F3BCC909ull, 0u, (uint32_t)NL};
            for (int b = 0; b < 600; b += NL) {
                lcg_refill(ls, jump, ns);
                if (b + lane < 600) pm.draws[b + lane] = ls.raw;
            }
        }
        wave_sync();
        // The side fails iff some valid line keeps fewer than 8 of the ns
        // inliers off it (ns - best < 8). The off-line count of a line only
        // grows with the points tested, so a line is settled once 8 points
        // lie off it: the wave tests the inliers in rank order 64 at a time
        // (one per lane, ballot + popcount) and leaves a line at the first
        // chunk that brings its off-line count to 8 -- the first chunk for
        // almost every line. Lines are prepared 64 at a time, lane L taking
        // line 64b + L, and read from that lane.
        const double2* xs = side ? t.x2 : t.x1;
        double2 q0 = make_double2(0.0, 0.0);
        if ((uint32_t)lane < ns) q0 = xs[inl_list[lane]];
        bool fail = false;
        for (int b0 = 0; b0 < 300 && !fail; b0 += NL) {
            double ax = 0.0, ay = 0.0, nx = 0.0, ny = 0.0;
            bool valid = false;
            const int it = b0 + lane;
            if (it < 300) {
                const uint32_t ri = pm.draws[2 * it], rj = pm.draws[2 * it + 1];
                const double2 a = xs[inl_list[ri]];
                const double2 bp = xs[inl_list[rj]];
                const double dx = bp.x - a.x, dy = bp.y - a.y;
                const double len = norm2(dx, dy);
                valid = !(ri == rj || isnan(len) || len <= 0.0);
                ax = a.x;
                ay = a.y;
                nx = -dy / len;
                ny = dx / len;
            }
            uint64_t vmask = __ballot(valid);
            while (vmask) {
                const int l = __builtin_ctzll(vmask);
                vmask &= vmask - 1;
                const double lax = readlane_d(ax, l), lay = readlane_d(ay, l);
                const double lnx = readlane_d(nx, l), lny = readlane_d(ny, l);
                uint32_t off = 0;
                for (uint32_t r0 = 0; r0 < ns; r0 += NL) {
                    const uint32_t r = r0 + (uint32_t)lane;
                    const double2 q = r0 == 0 ? q0 : (r < ns ? xs[inl_list[r]] : q0);
                    const double ex = q.x - lax, ey = q.y - lay;
                    const bool o = r < ns && !(fabs(ex * lnx + ey * lny) < cfg.threshold_px);
                    off += (uint32_t)__popcll(__ballot(o));
                    if (off >= 8) break;
                }
                if (off < 8) {
                    fail = true;
                    break;
                }
            }
        }
        if (lane == 0) sh.result = fail ? 0 : 1;
        wave_sync();
        if (!sh.result) {
            line_ok = false;
            break;
        }
    }
    lap(6);
    if (!line_ok) {
        if (lane == 0) *t.found = 0;
        flush_prof();
        return;
    }

    // --------------------------- select_two_view_model (matching.rs:425-492)
    RansacSel selo{};
    {
        PtsMask s{t.x1, t.x2, n, fin_mask};
        const double sigma_in = cfg.threshold_px / 3.0;
        bool has = ns >= 8 && isfinite(sigma_in) && sigma_in > 0.0;
        // F entries are finite (F is normalised); coordinates are finite.
        double sigma_hat = 0.0;
        if (has) {
            // median of the finite Sampson errors over the inliers, rank nf / 2;
            // the errors are computed once into LDS (or L2-resident scratch)
            double* err = n <= (uint32_t)(NL * kStageW) ? &coop().stage[0][0]
                                                        : reinterpret_cast<double*>(lo_masks);
            double f[9];
            load9(sh.f, f);
            uint32_t nfu = 0;
            for (uint32_t base = 0; base < n; base += NL) {
                const uint32_t i = base + lane;
                bool fin = false;
                if (i < n && s.take(i)) {
                    const double e = sampson_error(f, t.x1[i], t.x2[i]);
                    err[i] = e;
                    fin = isfinite(e);
                }
                nfu += (uint32_t)__popcll(__ballot(fin));
            }
            wave_sync();
            const int nf = (int)nfu;
            if (nf < 8) has = false;
            if (has) {
                const int want = nf / 2;
                double found_v = 0.0;
                int found_c = 0;
                for (uint32_t i = lane; i < n; i += NL) {
                    if (!s.take(i)) continue;
                    const double ei = err[i];
                    if (!isfinite(ei)) continue;
                    int below = 0, tie_before = 0;
                    for (uint32_t j = 0; j < n; ++j) {
                        if (!s.take(j)) continue;
                        const double ej = err[j];
                        if (!isfinite(ej)) continue;
                        below += ej < ei;
                        tie_before += (ej == ei && j < i);
                    }
                    if (below + tie_before == want) {
                        found_v = ei;
                        found_c = 1;
                    }
                }
                sh.redA[lane] = found_v;
                sh.lineBest[lane] = found_c;
                wave_sync();
                if (lane == 0) {
                    double m = 0.0;
                    for (int l = 0; l < NL; ++l)
                        if (sh.lineBest[l]) m = sh.redA[l];
                    sh.es_med = m;
                }
                wave_sync();
                sigma_hat = sqrt(sh.es_med / 0.455);
            }
        }
        if (has) {
            const double sigma = fmax(sigma_in, sigma_hat);
            const double s2 = sigma * sigma;
            double H[9], Hi[9];
            uint32_t hc = 0;
            const bool hok = fit_homography(t.x1, t.x2, n, fin_mask, inl_list, ns, 3.0 * sigma, H,
                                            Hi, &hc, pm.hom, fin_sel, list_buf);
            if (hok) {
                // GRIC sums over the inliers in index order (lanes 0 and 1)
                double f[9];
                load9(sh.f, f);
                double acc = 0.0;
                uint64_t parallax = 0;
                for (uint32_t base = 0; base < n; base += NL) {
                    const uint64_t m = bits64(fin_mask, base, n);
                    if (!m) continue;
                    bool par = false;
                    if ((m >> lane) & 1u) {
                        const uint32_t slot = (uint32_t)__popcll(m & lanes_below());
                        const double2 p = t.x1[base + lane], q = t.x2[base + lane];
                        const double ef = sampson_error(f, p, q) / s2;
                        double2 a, b;
                        const double e1 = htr(H, p, a) ? (a.x - q.x) * (a.x - q.x) + (a.y - q.y) * (a.y - q.y) : INFINITY;
                        const double e2 = htr(Hi, q, b) ? (b.x - p.x) * (b.x - p.x) + (b.y - p.y) * (b.y - p.y) : INFINITY;
                        const double eh = 0.25 * (e1 + e2) / s2;
                        coop().stage[slot][0] = fmin(ef, 2.0 * (4.0 - 3.0));
                        coop().stage[slot][1] = fmin(eh, 2.0 * (4.0 - 2.0));
                        par = eh > 13.8 && ef < 10.8;
                    }
                    parallax += (uint64_t)__popcll(__ballot(par));
                    wave_sync();
                    acc = stage_sum(lane & 1, (uint32_t)__popcll(m), acc);
                    wave_sync();
                }
                const double rho_f = readlane_d(acc, 0), rho_h = readlane_d(acc, 1);
                const double nf = (double)ns;
                const double l1 = log(4.0), l2 = log(4.0 * nf);
                selo.gric_f = rho_f + l1 * 3.0 * nf + l2 * 7.0;
                selo.gric_h = rho_h + l1 * 2.0 * nf + l2 * 8.0;
                selo.parallax = parallax;
                selo.parallax_needed = parallax_needed(ns);
                selo.model = (selo.gric_h < selo.gric_f && parallax < selo.parallax_needed) ? 1 : 0;
                selo.sigma_px = sigma;
                selo.has_selection = 1;
            }
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 9; ++i) t.F[i] = sh.f[i];
        *t.sel = selo;
        *t.found = 1;
    }
    uint32_t kept = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += NL) {
        const uint32_t i = i0 + lane;
        const bool keep = i < n && ((fin_mask[i >> 5] >> (i & 31)) & 1u);
        if (i < n) t.inliers[i] = keep ? 1u : 0u;
        if (t.ver_ij) {  // tracks.rs:56-69: the inlier matches in match order
            const unsigned long long m = __ballot(keep);
            if (keep) {
                const uint32_t k = kept + (uint32_t)__popcll(m & lanes_below());
                t.ver_ij[2 * k] = t.pair_ij[2 * i];
                t.ver_ij[2 * k + 1] = t.pair_ij[2 * i + 1];
            }
            kept += (uint32_t)__popcll(m);
        }
    }
    if (t.ver_len && lane == 0) *t.ver_len = kept;
    lap(7);
    flush_prof();
}

// Persistent workgroups over a work queue of pairs in `order` (costliest
// first): the first n_big pairs are large (a whole workgroup each, taken by
// workgroup from queue[0]), the rest small (a wave each, taken by wave from
// queue[1]). Every loop ends when its counter passes its range.
__global__ __launch_bounds__(kWaves * NL) __attribute__((amdgpu_waves_per_eu(SKH_RANSAC_WAVES_PER_EU)))
void ransac_kernel(const RansacTask* __restrict__ tasks, const uint32_t* __restrict__ order,
                   uint32_t n_tasks, uint32_t n_big, const RansacCfgDev cfg,
                   uint32_t* __restrict__ queue) {
    __shared__ uint32_t s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&queue[0], 1u);
        __syncthreads();
        const uint32_t it = s_item;
        __syncthreads();
        if (it >= n_big) break;
        run_pair(tasks[order ? order[it] : it], cfg, true);
        __syncthreads();
    }
    for (;;) {
        uint32_t it = 0;
        if (lane_id() == 0) it = atomicAdd(&queue[1], 1u);
        it = (uint32_t)__shfl((int)it, 0);
        if (it >= n_tasks - n_big) break;
        const uint32_t k = n_big + it;
        run_pair(tasks[order ? order[k] : k], cfg, false);
    }
}

__global__ void ransac_queue_reset_kernel(uint32_t* queue) {
    queue[0] = 0;
    queue[1] = 0;
}

// Cost estimate of each pair for the launch order: the first hypothesis
// batch of the pair (the same samples and solves as the real run), its best
// raw inlier count c and ratio w = c / n, and est = adaptive_iterations(w) *
// (64 + n) / 64 + 8 c: the hypotheses the run will draw times their cost in
// 64-point chunks, plus the pool refits and final stages, whose cost grows
// with the inlier count (a pair with many inliers draws few hypotheses but
// refits large selections). Pairs with a low inlier ratio run the full
// iteration budget whatever their size, so this orders the launch far better
// than the size alone; the weight 8 was measured best among 0-64.
__global__ __launch_bounds__(NL) __attribute__((amdgpu_waves_per_eu(SKH_RANSAC_WAVES_PER_EU)))
void ransac_probe_kernel(const RansacTask* __restrict__ tasks, const RansacCfgDev cfg,
                         uint32_t* __restrict__ est) {
    RansacTask t = tasks[blockIdx.x];
    if (t.n_dev) t.n = min(*t.n_dev, t.n);
    const int lane = lane_id();
    const uint32_t n = t.n;
    if (n < 8) {
        if (lane == 0) est[blockIdx.x] = 0;
        return;
    }
    const double th2 = cfg.threshold_px * cfg.threshold_px;
    const LcgLane jump = lcg_lane_jump();
    LcgStream smp{cfg.seed ^ 0x9E3779B97F4A7C15ull, 0u, (uint32_t)NL};
    PreBatch* pre = reinterpret_cast<PreBatch*>(t.scratch + (size_t)kWaves * kHandWords);
    lcg_samples<8>(smp, jump, n, kHypPerBatch, &pre->idx[0][0], 8);
    wave_sync();
    for (int round = 0; round < 3; ++round)
        group_solve(t.x1, t.x2, n, th2, nullptr, 0, pre, kGroups * round, kGroups);
    wave_sync();
    int best = lane < kHypPerBatch ? pre->cnt[lane] : 0;
#pragma unroll
    for (int off = 1; off < NL; off <<= 1) {
        const int o = __shfl_xor(best, off);
        best = best > o ? best : o;
    }
    if (lane == 0) {
        const double w = best > 0 ? (double)best / (double)n : 0.0;
        const uint64_t it = ipow_adaptive(w, 8, cfg.confidence, cfg.max_iters);
        const double c = (double)it * (64.0 + (double)n) / 64.0 + 8.0 * (double)best;
        est[blockIdx.x] = c < 4.0e9 ? (uint32_t)c : 4000000000u;
    }
}

}  // namespace

// Pair p's rank when the pairs are ordered by descending cost estimate
// (ties by index): order[rank] = p. The costliest problems start first, so
// the longest workgroups do not trail at the end of the launch.
// Pairs in descending estimated cost, ties by index: pair i's rank is the number
// of pairs ahead of it. A workgroup ranks 64 pairs; its four waves count over
// interleaved quarters of the list (four independent counters per lane), then
// the partial counts are added.
__global__ __launch_bounds__(256) void ransac_order_kernel(const uint32_t* __restrict__ count,
                                                           uint32_t P, uint32_t* __restrict__ order) {
    __shared__ uint32_t part[4][64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * 64u + lane;
    const uint32_t ki = i < P ? count[i] : 0u;
    uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    // groups of four consecutive pairs, group g to wave g mod 4
    for (uint32_t j = 4u * wave; j < P; j += 16u) {
        if (j + 3u < P) {
            const uint32_t k0 = count[j], k1 = count[j + 1], k2 = count[j + 2], k3 = count[j + 3];
            r0 += (k0 > ki || (k0 == ki && j < i)) ? 1u : 0u;
            r1 += (k1 > ki || (k1 == ki && j + 1u < i)) ? 1u : 0u;
            r2 += (k2 > ki || (k2 == ki && j + 2u < i)) ? 1u : 0u;
            r3 += (k3 > ki || (k3 == ki && j + 3u < i)) ? 1u : 0u;
        } else {
            for (uint32_t jj = j; jj < P; ++jj) {
                const uint32_t kj = count[jj];
                r0 += (kj > ki || (kj == ki && jj < i)) ? 1u : 0u;
            }
        }
    }
    part[wave][lane] = r0 + r1 + r2 + r3;
    __syncthreads();
    if (wave == 0 && i < P)
        order[part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]] = i;
}

void launch_ransac(const RansacTask* tasks, uint32_t n_tasks, const RansacCfgDev& cfg,
                   hipStream_t s, uint32_t* queue, uint32_t* est, uint32_t* order) {
    if (!n_tasks) return;
    if (est && order) {
        hipLaunchKernelGGL(ransac_probe_kernel, dim3(n_tasks), dim3(NL), 0, s, tasks, cfg, est);
        hipLaunchKernelGGL(ransac_order_kernel, dim3((n_tasks + 63) / 64), dim3(256), 0, s,
                           (const uint32_t*)est, n_tasks, order);
    } else {
        order = nullptr;
    }
    hipLaunchKernelGGL(ransac_queue_reset_kernel, dim3(1), dim3(1), 0, s, queue);
    // the costliest 1/32 of the pairs (a single pair: itself) get a whole workgroup
    const uint32_t n_big = order ? std::min(n_tasks, std::max(1u, n_tasks / 32)) : 1u;
    // resident workgroups: two per CU (64 KB LDS and 256 VGPRs per lane each)
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
    }
    const uint32_t need = n_big + (n_tasks - n_big + kWaves - 1) / kWaves;
    const uint32_t grid = std::min(need, 2u * (uint32_t)cus);
    hipLaunchKernelGGL(ransac_kernel, dim3(grid), dim3(kWaves * NL), 0, s, tasks,
                       (const uint32_t*)order, n_tasks, n_big, cfg, queue);
}

uint32_t ransac_scratch_words(uint32_t n) {
    // kWaves sets of LO hand-off records, the batch hypotheses, then the
    // bitsets and kWaves index lists; a multiple of 64 words keeps every
    // pair's scratch 256-byte aligned
    const uint32_t w = kWaves * kHandWords + kPreWords + kMaskSets * words(n);
    return (w + 63u) & ~63u;
}

}  // namespace skh
