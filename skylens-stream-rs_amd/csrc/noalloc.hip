// The allocation-free form of skh_ratio_match_dev, for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation). Every device buffer the pass needs
// beyond the caller's inputs and outputs -- quantised rows, row and column
// state, column partials, the exact re-score queue, task records -- lives in
// one file-scope device array, `scratch`, of fixed size, sized for the
// reference's feature cap of 8192 per view (features.rs:285 max_features).
// Each GPU holds its own copy of the array until the process ends; every call
// on a GPU shares it, so calls that overlap on one GPU (several contexts, or
// streams, at once) race, and each call rewrites what it reads.
#include <mutex>

#include "skh_ctx.hpp"
#include "skh_lowp.hpp"
#include "skh_single.hpp"
#include "skylens_hip.h"

namespace skh {
namespace {

constexpr size_t kRows = 8192;  // rows per side
constexpr size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Region sizes, each for the largest need over the precisions and tiles:
// bf16 rows take 256 B (e4m3 rows 128 B), the e4m3 metadata 8 B per row (the
// bf16 norm 4 B), column partials one float2 per (row block, column) with row
// blocks of at least 128 rows, the exact pass's row blocks 64 rows.
constexpr size_t kQ = align_up(2 * kDescLen * kRows);
constexpr size_t kMeta = align_up(8 * kRows);
constexpr size_t kRow4 = align_up(4 * kRows);
constexpr size_t kCol8 = align_up(8 * kRows);
constexpr size_t kParts = align_up(8 * kRows * (kRows / 128));
constexpr size_t kQueue = align_up(16 * kRows);
constexpr size_t kChunks = align_up(8 * chunk_capacity(kRows, 1));
constexpr size_t kBlocks = align_up(sizeof(RowBlock) * (kRows / 64));
constexpr size_t kSmall = 256;
constexpr size_t kScratchBytes = 2 * kQ + 2 * kMeta + 6 * kRow4 + 3 * kCol8 + kParts + 2 * kQueue +
                                 kChunks + kBlocks + 5 * kSmall;
static_assert(sizeof(LowpTask) <= kSmall && sizeof(PairTask) <= kSmall &&
                  2 * sizeof(ViewStats) <= kSmall,
              "task records must fit their regions");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Carves the regions out of the array's device address `base`.
SingleBufs carve(uint8_t* base) {
    size_t off = 0;
    auto take = [&](size_t n) {
        uint8_t* p = base + off;
        off += n;
        return p;
    };
    SingleBufs B{};
    uint8_t* qa = take(kQ);
    uint8_t* qb = take(kQ);
    B.q_a = reinterpret_cast<uint16_t*>(qa);
    B.q_b = reinterpret_cast<uint16_t*>(qb);
    B.q8_a = qa;
    B.q8_b = qb;
    uint8_t* ma = take(kMeta);
    uint8_t* mb = take(kMeta);
    B.nrm_a = reinterpret_cast<uint32_t*>(ma);
    B.nrm_b = reinterpret_cast<uint32_t*>(mb);
    B.meta_a = reinterpret_cast<Fp8Meta*>(ma);
    B.meta_b = reinterpret_cast<Fp8Meta*>(mb);
    B.row_k1 = reinterpret_cast<float*>(take(kRow4));
    B.row_j1 = reinterpret_cast<uint32_t*>(take(kRow4));
    B.row_k2 = reinterpret_cast<float*>(take(kRow4));
    B.row_match = reinterpret_cast<uint32_t*>(take(kRow4));
    B.row_d1 = reinterpret_cast<float*>(take(kRow4));
    B.row_d2 = reinterpret_cast<float*>(take(kRow4));
    B.col_top2 = reinterpret_cast<float2*>(take(kCol8));
    B.col_key = reinterpret_cast<unsigned long long*>(take(kCol8));
    (void)take(kCol8);  // reserve
    B.col_part = reinterpret_cast<float2*>(take(kParts));
    B.queue = reinterpret_cast<uint4*>(take(kQueue));
    B.pre = reinterpret_cast<uint4*>(take(kQueue));
    B.chunks = reinterpret_cast<uint2*>(take(kChunks));
    B.blocks = reinterpret_cast<RowBlock*>(take(kBlocks));
    B.stats = reinterpret_cast<ViewStats*>(take(kSmall));
    B.ltask = reinterpret_cast<LowpTask*>(take(kSmall));
    B.task = reinterpret_cast<PairTask*>(take(kSmall));
    B.queue_len = reinterpret_cast<uint32_t*>(take(kSmall));
    B.out_base = reinterpret_cast<uint64_t*>(take(kSmall));
    return B;
}

}  // namespace

// About 9.2 MB (8.75 MiB) per GPU.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// The device address of `scratch` on `device` (which must be current), looked
// up once per device. In host code the array's name stands for no device
// memory, so everything that touches the array goes through this address.
int noalloc_scratch_address(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return SKH_ENODEV;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return SKH_EHIP;
        }
        g_ptr[device] = p;
    }
    *out = g_ptr[device];
    return SKH_OK;
}

size_t noalloc_max_rows() { return kRows; }

}  // namespace skh

using namespace skh;

extern "C" int skh_ratio_match_dev_noalloc(skh_ctx* ctx, const float* d_a, size_t na,
                                           const float* d_b, size_t nb, float ratio, int mutual,
                                           int precision, uint32_t* d_out_ij,
                                           uint32_t* d_out_len) {
    if (!ctx) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if ((!d_a && na) || (!d_b && nb) || !d_out_len || (!d_out_ij && na) ||
        !skh_precision_supported(precision))
        return ctx->fail(SKH_EINVAL, "skh_ratio_match_dev_noalloc: invalid argument");
    if (na > kRows || nb > kRows)
        return ctx->fail(SKH_EINVAL,
                         "skh_ratio_match_dev_noalloc: %zu x %zu rows, at most %zu per side", na,
                         nb, kRows);
    if (hipSetDevice(ctx->device) != hipSuccess) return SKH_ENODEV;
    if (!ctx->noalloc_scratch)
        return ctx->fail(SKH_EHIP, "skh_ratio_match_dev_noalloc: scratch address unavailable");
    const SingleBufs B = carve(static_cast<uint8_t*>(ctx->noalloc_scratch));
    const uint32_t a_rows = (uint32_t)na, b_rows = (uint32_t)nb;
    if (precision == SKH_PREC_FP32) {
        enqueue_exact_single(B, d_a, a_rows, d_b, b_rows, ratio, mutual != 0, d_out_ij, d_out_len,
                             ctx->stream);
    } else if (na == 0 || nb < 2) {
        launch_fill_u32(d_out_len, 1, 0u, ctx->stream);  // matching.rs:206: no second neighbour
    } else {
        enqueue_lowp_single(B, d_a, a_rows, d_b, b_rows, ratio, mutual != 0,
                            precision == SKH_PREC_FP8, ctx->pick_tile(precision, na, nb),
                            ctx->num_cu, d_out_ij, d_out_len, ctx->stream);
    }
    return ctx->hip(hipGetLastError(), "skh_ratio_match_dev_noalloc: launch");
}
