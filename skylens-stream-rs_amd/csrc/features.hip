// Feature detector on the GPU: detect_and_describe (crates/core/src/features.rs:594-781).
//
// Per image (and, for small images, again on the 2x bilinear upsample):
//   blur_h / blur_v      separable Gaussian of gaussian_blur (features.rs:214-256),
//                        taps computed on the host with expf exactly as the
//                        reference does, each output summed in tap order; the
//                        vertical pass also writes the DoG layer (gauss[i] -
//                        gauss[i-1], features.rs:681-693)
//   grad_kernel          gradient table of layers 1..s (features.rs:340-352)
//   extrema_kernel       26-neighbour extrema, refine_extremum, contrast and
//                        edge tests (features.rs:717-758), one candidate record
//                        per surviving scan position
//   radix sort           candidates into the reference's emission order
//                        (octave, layer, row, column)
//   orient_kernel        dominant orientations (features.rs:361-413), one key
//                        per (candidate, angle): |response| descending, then
//                        emission order -- the stable sort of features.rs:778
//   radix sort           features by that key; the first max_features are kept
//   describe_kernel      the 128-d descriptor (features.rs:521-592), computed
//                        only for the kept features
//   dedup / select       the upsampled image's features that do not duplicate
//                        an original one (features.rs:606-646)
//
// Arithmetic is f32 in the reference's order (no FMA contraction); the f32
// transcendentals evaluated per pixel or keypoint are the correctly rounded
// ones of mathf_dev.hpp.
#pragma clang fp contract(off)

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mathf_dev.hpp"
#include "skh_ctx.hpp"
#include "skylens_hip.h"

namespace skh {
namespace {

constexpr float kTau = 6.28318548202514648438f;     // f32 TAU
constexpr float kSqrt2f = 1.41421353816986083984f;  // f32 SQRT_2
constexpr int kMaxR = 64;                           // largest Gaussian radius (sigma <= 21)
constexpr int kMaxLayers = 16;                      // s + 2 DoG layers, s <= 14
constexpr int kMaxOctaves = 16;
constexpr int kMaxAngles = 18;  // strict local maxima of a 36-bin circular histogram

struct Taps {
    int r;
    float k[2 * kMaxR + 1];
};

// f32 as isize (saturating, NaN -> 0); only small values occur.
__device__ __forceinline__ int to_int(float v) {
    if (v != v) return 0;
    v = fminf(fmaxf(v, -2.0e9f), 2.0e9f);
    return (int)v;
}
__device__ __forceinline__ float rem_euclid(float a, float b) {
    const float r = fmodf(a, b);
    return r < 0.0f ? r + fabsf(b) : r;
}

// Horizontal pass: each workgroup stages its row segment plus the clamped
// halo in LDS, each thread sums its taps in order.
__global__ __launch_bounds__(256) void blur_h_kernel(const float* __restrict__ src,
                                                     float* __restrict__ dst, int w, Taps t) {
    __shared__ float row[256 + 2 * kMaxR];
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * 256;
    const float* s = src + (size_t)y * w;
    for (int i = threadIdx.x; i < 256 + 2 * t.r; i += 256) {
        const int xx = min(max(x0 + i - t.r, 0), w - 1);
        row[i] = s[xx];
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= w) return;
    float acc = 0.0f;
    for (int j = 0; j <= 2 * t.r; ++j) acc = acc + t.k[j] * row[threadIdx.x + j];
    dst[(size_t)y * w + x] = acc;
}

// Vertical pass: out = sum over taps in order of k[j] * tmp[clamp(y + j - r)];
// with `prev` (the blur's own input) also the DoG layer out - prev.
__global__ __launch_bounds__(256) void blur_v_kernel(const float* __restrict__ tmp,
                                                     float* __restrict__ dst, int w, int h, Taps t,
                                                     const float* __restrict__ prev,
                                                     float* __restrict__ dog) {
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    float acc = 0.0f;
    for (int j = 0; j <= 2 * t.r; ++j) {
        const int yy = min(max(y + j - t.r, 0), h - 1);
        acc = acc + t.k[j] * tmp[(size_t)yy * w + x];
    }
    const size_t i = (size_t)y * w + x;
    dst[i] = acc;
    if (dog) dog[i] = acc - prev[i];
}

// Gradient table (magnitude, atan2) of the interior; the border is never read.
__global__ __launch_bounds__(256) void grad_kernel(const float* __restrict__ g,
                                                   float2* __restrict__ out, int w, int h) {
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    float2 v = make_float2(0.0f, 0.0f);
    if (x >= 1 && x + 1 < w && y >= 1 && y + 1 < h) {
        const size_t i = (size_t)y * w + x;
        const float gx = g[i + 1] - g[i - 1];
        const float gy = g[i + w] - g[i - w];
        v.x = sqrtf(gx * gx + gy * gy);
        v.y = fm::atan2f_cr(gy, gx);
    }
    out[(size_t)y * w + x] = v;
}

__global__ __launch_bounds__(256) void downsample_kernel(const float* __restrict__ src, int w,
                                                         float* __restrict__ dst, int W, int H) {
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W || y >= H) return;
    dst[(size_t)y * W + x] = src[(size_t)(2 * y) * w + 2 * x];
}

// GrayImage::upsample2 (features.rs:55-81).
__global__ __launch_bounds__(256) void upsample_kernel(const float* __restrict__ src, int w, int h,
                                                       float* __restrict__ dst) {
    const int W = 2 * w;
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W) return;
    const float sy = ((float)y + 0.5f) / 2.0f - 0.5f;
    const float y0 = floorf(sy);
    const float fy = sy - y0;
    const int ya = min(max(to_int(y0), 0), h - 1), yb = min(max(to_int(y0) + 1, 0), h - 1);
    const float sx = ((float)x + 0.5f) / 2.0f - 0.5f;
    const float x0 = floorf(sx);
    const float fx = sx - x0;
    const int xa = min(max(to_int(x0), 0), w - 1), xb = min(max(to_int(x0) + 1, 0), w - 1);
    const float top = src[(size_t)ya * w + xa] * (1.0f - fx) + src[(size_t)ya * w + xb] * fx;
    const float bot = src[(size_t)yb * w + xa] * (1.0f - fx) + src[(size_t)yb * w + xb] * fx;
    dst[(size_t)y * W + x] = top * (1.0f - fy) + bot * fy;
}

// One surviving extremum, at its scan position (the emission order) with the
// refined sample, sub-pixel position and scale coordinate.
struct Cand {
    uint64_t key;  // octave << 56 | layer << 48 | y << 24 | x (scan position)
    float x, y, s, value;
    uint32_t xi, yi, layer, octave;
    float sig;  // sigma0 * kstep^s, filled by orient_kernel
    uint32_t pad;
};

struct DogSet {
    const float* d[kMaxLayers];
};

// refine_extremum (features.rs:430-485); the 3x3 inverse is the cofactor form.
__device__ bool refine(const DogSet& D, int w, int h, int x, int y, int l, int s, Cand& c) {
    for (int it = 0; it < 5; ++it) {
        const float* dm = D.d[l - 1];
        const float* d0 = D.d[l];
        const float* dp = D.d[l + 1];
        auto at = [&](const float* d, int xx, int yy) { return d[(size_t)yy * w + xx]; };
        const float v = at(d0, x, y);
        const float g0 = 0.5f * (at(d0, x + 1, y) - at(d0, x - 1, y));
        const float g1 = 0.5f * (at(d0, x, y + 1) - at(d0, x, y - 1));
        const float g2 = 0.5f * (at(dp, x, y) - at(dm, x, y));
        const float hxx = at(d0, x + 1, y) + at(d0, x - 1, y) - 2.0f * v;
        const float hyy = at(d0, x, y + 1) + at(d0, x, y - 1) - 2.0f * v;
        const float hss = at(dp, x, y) + at(dm, x, y) - 2.0f * v;
        const float hxy = 0.25f * (at(d0, x + 1, y + 1) - at(d0, x - 1, y + 1) -
                                   at(d0, x + 1, y - 1) + at(d0, x - 1, y - 1));
        const float hxs =
            0.25f * (at(dp, x + 1, y) - at(dp, x - 1, y) - at(dm, x + 1, y) + at(dm, x - 1, y));
        const float hys =
            0.25f * (at(dp, x, y + 1) - at(dp, x, y - 1) - at(dm, x, y + 1) + at(dm, x, y - 1));
        const float m11 = hxx, m12 = hxy, m13 = hxs, m21 = hxy, m22 = hyy, m23 = hys, m31 = hxs,
                    m32 = hys, m33 = hss;
        const float c11 = m22 * m33 - m32 * m23;
        const float c21 = m21 * m33 - m31 * m23;
        const float c31 = m21 * m32 - m31 * m22;
        const float det = m11 * c11 - m12 * c21 + m13 * c31;
        if (det == 0.0f) return false;
        const float i0 = c11 / det, i1 = (m13 * m32 - m33 * m12) / det,
                    i2 = (m12 * m23 - m22 * m13) / det;
        const float i3 = -c21 / det, i4 = (m11 * m33 - m31 * m13) / det,
                    i5 = (m13 * m21 - m23 * m11) / det;
        const float i6 = c31 / det, i7 = (m12 * m31 - m32 * m11) / det,
                    i8 = (m11 * m22 - m21 * m12) / det;
        const float d0x = -(i0 * g0 + i1 * g1 + i2 * g2);
        const float d0y = -(i3 * g0 + i4 * g1 + i5 * g2);
        const float d0s = -(i6 * g0 + i7 * g1 + i8 * g2);
        if (fabsf(d0x) <= 0.5f && fabsf(d0y) <= 0.5f && fabsf(d0s) <= 0.5f) {
            c.xi = (uint32_t)x;
            c.yi = (uint32_t)y;
            c.layer = (uint32_t)l;
            c.x = (float)x + d0x;
            c.y = (float)y + d0y;
            c.s = (float)l + d0s;
            c.value = v + 0.5f * (g0 * d0x + g1 * d0y + g2 * d0s);
            return true;
        }
        if (!isfinite(d0x) || !isfinite(d0y) || !isfinite(d0s)) return false;
        const int nx = x + to_int(roundf(d0x));
        const int ny = y + to_int(roundf(d0y));
        const int nl = l + to_int(roundf(d0s));
        if (nx < 1 || ny < 1 || nl < 1 || nx >= w - 1 || ny >= h - 1 || nl > s) return false;
        x = nx;
        y = ny;
        l = nl;
    }
    return false;
}

// One thread per interior scan position (x, y) of layer l = 1 + blockIdx.z.
__global__ __launch_bounds__(256) void extrema_kernel(DogSet D, int w, int h, int s, int octave,
                                                      float contrast, float edge_thr,
                                                      Cand* __restrict__ out,
                                                      uint32_t* __restrict__ count, uint32_t cap) {
    const int x = 1 + blockIdx.x * 256 + threadIdx.x;
    const int y = 1 + blockIdx.y;
    const int l = 1 + blockIdx.z;
    if (x >= w - 1) return;
    const float* d1 = D.d[l];
    const float v = d1[(size_t)y * w + x];
    if (fabsf(v) < contrast) return;
    bool is_max = true, is_min = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float* d = D.d[l - 1 + k];
#pragma unroll
        for (int yy = -1; yy <= 1; ++yy)
#pragma unroll
            for (int xx = -1; xx <= 1; ++xx) {
                if (k == 1 && xx == 0 && yy == 0) continue;
                const float u = d[(size_t)(y + yy) * w + (x + xx)];
                is_max &= v > u;
                is_min &= v < u;
            }
    }
    if (!(is_max || is_min)) return;
    Cand c;
    if (!refine(D, w, h, x, y, l, s, c)) return;
    if (fabsf(c.value) < contrast) return;
    const float* d = D.d[c.layer];
    const int xi = (int)c.xi, yi = (int)c.yi;
    auto at = [&](int xx, int yy) { return d[(size_t)yy * w + xx]; };
    const float cc = at(xi, yi);
    const float dxx = at(xi + 1, yi) + at(xi - 1, yi) - 2.0f * cc;
    const float dyy = at(xi, yi + 1) + at(xi, yi - 1) - 2.0f * cc;
    const float dxy = 0.25f * (at(xi + 1, yi + 1) - at(xi - 1, yi + 1) - at(xi + 1, yi - 1) +
                               at(xi - 1, yi - 1));
    const float tr = dxx + dyy;
    const float det = dxx * dyy - dxy * dxy;
    if (det <= 0.0f || tr * tr / det >= edge_thr) return;
    c.key = ((uint64_t)octave << 56) | ((uint64_t)l << 48) | ((uint64_t)y << 24) | (uint64_t)x;
    c.octave = (uint32_t)octave;
    c.sig = 0.0f;
    c.pad = 0u;
    const uint32_t slot = atomicAdd(count, 1u);
    if (slot < cap) out[slot] = c;
}

// Grad tables of every (octave, layer) of one pyramid.
struct GradSet {
    const float2* g[kMaxOctaves][kMaxLayers];
    int w[kMaxOctaves], h[kMaxOctaves];
};

// Dominant orientations (features.rs:361-413) of candidate order[c], one
// wave per candidate. The window's pixels are taken 64 at a time in the
// reference's raster order: each lane stages one pixel's bin and weighted
// magnitude in LDS, then lane b < 36 (the owner of bin b) adds the staged
// values of its bin in pixel order, which is the reference's summation order
// for every bin. Writes one sort key per angle: (~|response| bits) << 32 |
// c << 5 | angle index, and the angle's bits as the value.
__global__ __launch_bounds__(64) void orient_kernel(const GradSet* __restrict__ G,
                                                    Cand* __restrict__ cands,
                                                    const uint32_t* __restrict__ order,
                                                    uint32_t n, float sigma0, float kstep,
                                                    uint64_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals,
                                                    uint32_t* __restrict__ nfeat) {
    constexpr int NB = 36;
    __shared__ int sbin[64];
    __shared__ float sval[64];
    __shared__ float hbuf[2][NB];
    const uint32_t c = blockIdx.x;
    const int lane = threadIdx.x;
    if (c >= n) return;
    Cand* cd = &cands[order[c]];
    const float sig = sigma0 * fm::powf_pos(kstep, cd->s);
    if (lane == 0) cd->sig = sig;
    const int w = G->w[cd->octave], h = G->h[cd->octave];
    const float2* g = G->g[cd->octave][cd->layer];
    const float x = cd->x, y = cd->y;
    const float sw = 1.5f * sig;
    const int r = to_int(roundf(3.0f * sw));
    const int xi = to_int(roundf(x)), yi = to_int(roundf(y));
    const int side = 2 * r + 1, npx = side * side;
    float hv = 0.0f;  // bin `lane` (lanes < 36)
    for (int p0 = 0; p0 < npx; p0 += 64) {
        const int p = p0 + lane;
        int bin = -1;
        float val = 0.0f;
        if (p < npx) {
            const int px = xi + (p % side - r), py = yi + (p / side - r);
            if (px >= 1 && py >= 1 && px < w - 1 && py < h - 1) {
                const float2 mg = g[(size_t)py * w + px];
                const float fx = (float)px - x, fy = (float)py - y;
                const float wgt = fm::expf_cr(-(fx * fx + fy * fy) / (2.0f * sw * sw));
                const float ang = rem_euclid(mg.y, kTau);
                bin = to_int(roundf(ang / kTau * (float)NB)) % NB;
                val = wgt * mg.x;
            }
        }
        const uint64_t valid = __ballot(bin >= 0);
        if (valid == 0) continue;
        const int cnt = __popcll(valid);
        if (bin >= 0) {
            const int slot = __popcll(valid & ((1ull << lane) - 1ull));
            sbin[slot] = bin;
            sval[slot] = val;
        }
        __syncthreads();
        // +0.0 for the other bins leaves the (never -0.0) sums unchanged
#pragma unroll 4
        for (int j = 0; j < cnt; ++j) hv = hv + (sbin[j] == lane ? sval[j] : 0.0f);
        __syncthreads();
    }
    // six passes of [1, 1, 1] / 3, then the peaks
    for (int pass = 0; pass < 6; ++pass) {
        if (lane < NB) hbuf[0][lane] = hv;
        __syncthreads();
        if (lane < NB)
            hv = (hbuf[0][(lane + NB - 1) % NB] + hbuf[0][lane] + hbuf[0][(lane + 1) % NB]) / 3.0f;
        __syncthreads();
    }
    if (lane < NB) hbuf[1][lane] = hv;
    __syncthreads();
    float mx = 0.0f;
    for (int i = 0; i < NB; ++i) mx = fmaxf(mx, hbuf[1][i]);
    if (mx <= 0.0f) return;
    bool peak = false;
    float angle = 0.0f;
    if (lane < NB) {
        const float l = hbuf[1][(lane + NB - 1) % NB], cc = hv, rr = hbuf[1][(lane + 1) % NB];
        if (cc >= 0.8f * mx && cc > l && cc > rr) {
            peak = true;
            const float off = 0.5f * (l - rr) / (l - 2.0f * cc + rr);
            angle = rem_euclid(((float)lane + off) / (float)NB * kTau, kTau);
        }
    }
    const uint64_t pm = __ballot(peak);
    if (pm == 0) return;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(nfeat, (uint32_t)__popcll(pm));
    base = __shfl(base, 0);
    if (peak) {
        const uint32_t a = (uint32_t)__popcll(pm & ((1ull << lane) - 1ull));
        const float resp = fabsf(cd->value);
        const uint64_t hi = (uint64_t)(0xffffffffu - __float_as_uint(resp)) << 32;
        keys[base + a] = hi | ((uint64_t)c << 5) | a;
        vals[base + a] = __float_as_uint(angle);
    }
}

// Output slot of a kept feature, its keypoint in output coordinates.
struct DescJob {
    const uint64_t* keys;    // sorted feature keys
    const uint32_t* angles;  // sorted angle bits
    const uint32_t* order;   // sorted candidate order
    const Cand* cands;
    const GradSet* G;
    const uint32_t* sel;  // feature index per output (NULL: output k = feature k)
    uint32_t n;           // outputs
    uint32_t out0;        // first output slot
    int upscaled;         // features of the 2x image: x / 2 - 0.25, sigma / 2
};

__device__ __forceinline__ void feature_kp(const DescJob& J, uint32_t f, skh_keypoint& kp,
                                           const Cand*& cd, float& angle) {
    const uint64_t key = J.keys[f];
    const uint32_t c = (uint32_t)((key >> 5) & 0x7ffffffu);
    cd = &J.cands[J.order[c]];
    angle = __uint_as_float(J.angles[f]);
    const float scale = (float)(1u << cd->octave);
    kp.x = cd->x * scale;
    kp.y = cd->y * scale;
    kp.sigma = cd->sig * scale;
    kp.response = cd->value;
    kp.angle = angle;
    if (J.upscaled) {
        kp.x = kp.x / 2.0f - 0.25f;
        kp.y = kp.y / 2.0f - 0.25f;
        kp.sigma = kp.sigma / 2.0f;
    }
}

// describe_with (features.rs:521-592), one wave per kept feature. The
// window's pixels are taken 64 at a time in the reference's raster order:
// each lane computes one pixel's weighted magnitude m and its trilinear cell
// (v0, u0, o0 and the fractions) and stages them in LDS; then lane l, the
// owner of bins 2l and 2l + 1 (spatial cell l / 4, orientations 2 (l % 4) and
// 2 (l % 4) + 1), adds each staged pixel's contribution m * wv * wu * wo to
// its bins, in pixel order -- every bin's sum has the reference's operands in
// the reference's order.
__global__ __launch_bounds__(64) void describe_kernel(DescJob J, skh_keypoint* __restrict__ kps,
                                                      float* __restrict__ desc) {
    constexpr int NC = 4, NO = 8;
    __shared__ float4 sent[64];  // m, du, dv, dob of the staged pixels
    __shared__ int spk[64];      // 1 | (v0 + 1) << 1 | (u0 + 1) << 4 | o0 << 7
    __shared__ float hist[128];
    __shared__ float snorm;
    const uint32_t k = blockIdx.x;
    const int lane = threadIdx.x;
    if (k >= J.n) return;
    const uint32_t f = J.sel ? J.sel[k] : k;
    skh_keypoint kp;
    const Cand* cd;
    float angle;
    feature_kp(J, f, kp, cd, angle);
    const uint32_t o = J.out0 + k;
    if (kps && lane == 0) kps[o] = kp;
    if (!desc) return;
    const int w = J.G->w[cd->octave], h = J.G->h[cd->octave];
    const float2* g = J.G->g[cd->octave][cd->layer];
    const float x = cd->x, y = cd->y, sigma = cd->sig;
    const float cell = 3.0f * sigma;
    const int r = to_int(roundf(cell * kSqrt2f * ((float)NC + 1.0f) * 0.5f));
    float sn, cs;
    fm::sincosf_cr(angle, &sn, &cs);
    const int xi = to_int(roundf(x)), yi = to_int(roundf(y));
    const int side = 2 * r + 1, npx = side * side;
    const int my_cell = lane >> 2, cv = my_cell >> 2, cu = my_cell & 3;
    const int oa = 2 * (lane & 3), ob = oa + 1;
    float ha = 0.0f, hb = 0.0f;  // bins 2 lane, 2 lane + 1
    for (int p0 = 0; p0 < npx; p0 += 64) {
        const int p = p0 + lane;
        int pk = 0;
        float m = 0.0f, du = 0.0f, dv = 0.0f, dob = 0.0f;
        if (p < npx) {
            const int px = xi + (p % side - r), py = yi + (p / side - r);
            if (px >= 1 && py >= 1 && px < w - 1 && py < h - 1) {
                const float fx = (float)px - x, fy = (float)py - y;
                const float u = (cs * fx + sn * fy) / cell;
                const float v = (-sn * fx + cs * fy) / cell;
                const float bu = u + (float)NC / 2.0f - 0.5f, bv = v + (float)NC / 2.0f - 0.5f;
                if (!(bu <= -1.0f || bv <= -1.0f || bu >= (float)NC || bv >= (float)NC)) {
                    const float2 mg = g[(size_t)py * w + px];
                    const float ori = rem_euclid(mg.y - angle, kTau);
                    const float bo = ori / kTau * (float)NO;
                    const float wgt = fm::expf_cr(-(u * u + v * v) / 8.0f);  // 2 (NC/2)^2 = 8
                    m = mg.x * wgt;
                    const float u0 = floorf(bu), v0 = floorf(bv), o0 = floorf(bo);
                    du = bu - u0;
                    dv = bv - v0;
                    dob = bo - o0;
                    pk = 1 | ((to_int(v0) + 1) << 1) | ((to_int(u0) + 1) << 4) | (to_int(o0) << 7);
                }
            }
        }
        // stage the contributing pixels only, compacted in pixel order
        const uint64_t valid = __ballot(pk != 0);
        if (valid == 0) continue;
        const int cnt = __popcll(valid);
        if (pk) {
            const int slot = __popcll(valid & ((1ull << lane) - 1ull));
            sent[slot] = make_float4(m, du, dv, dob);
            spk[slot] = pk;
        }
        __syncthreads();
        // Every entry is added to both bins, as the product of the reference's
        // weights when the bin is one of the pixel's eight, else as +0.0: the
        // partial sums are never -0.0 (every term is >= +0), so adding +0.0
        // leaves them bit for bit unchanged.
#pragma unroll 4
        for (int j = 0; j < cnt; ++j) {
            const int q = spk[j];
            const float4 e = sent[j];
            const int a = cv - (((q >> 1) & 7) - 1);  // 0: row v0 (weight 1 - dv), 1: v0 + 1
            const int b = cu - (((q >> 4) & 7) - 1);
            const bool cellin = (unsigned)a <= 1u && (unsigned)b <= 1u;
            const int io0 = (q >> 7) & 7, io1 = ((q >> 7) + 1) & 7;
            const float wv = a ? e.z : 1.0f - e.z;
            const float wu = b ? e.y : 1.0f - e.y;
            const float mvu = e.x * wv * wu;
            const float woa = oa == io0 ? 1.0f - e.w : e.w;
            const float wob = ob == io0 ? 1.0f - e.w : e.w;
            const float ca = mvu * woa, cb = mvu * wob;
            ha = ha + ((cellin && (oa == io0 || oa == io1)) ? ca : 0.0f);
            hb = hb + ((cellin && (ob == io0 || ob == io1)) ? cb : 0.0f);
        }
        __syncthreads();
    }
    // normalise (sum of squares in bin order), clip at 0.2, normalise again
    hist[2 * lane] = ha;
    hist[2 * lane + 1] = hb;
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
        if (lane == 0) {
            float s = -0.0f;
            for (int i = 0; i < 128; ++i) s = s + hist[i] * hist[i];
            snorm = sqrtf(s);
        }
        __syncthreads();
        const float nrm = snorm;
        float va = hist[2 * lane], vb = hist[2 * lane + 1];
        if (nrm > 0.0f) {
            va = va / nrm;
            vb = vb / nrm;
        }
        if (pass == 0) {
            va = fminf(va, 0.2f);
            vb = fminf(vb, 0.2f);
        }
        __syncthreads();
        hist[2 * lane] = va;
        hist[2 * lane + 1] = vb;
        __syncthreads();
    }
    float2* out = reinterpret_cast<float2*>(desc + (size_t)o * 128);
    out[lane] = make_float2(hist[2 * lane], hist[2 * lane + 1]);
}

// Extra (upsampled) feature e duplicates an original one when it lies within
// one pixel and its scale ratio is at most 1.5 (features.rs:627-638).
__global__ __launch_bounds__(256) void dedup_kernel(DescJob J, const skh_keypoint* __restrict__ orig,
                                                    uint32_t n_orig, uint32_t* __restrict__ keep) {
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= J.n) return;
    skh_keypoint kp;
    const Cand* cd;
    float angle;
    feature_kp(J, e, kp, cd, angle);
    uint32_t dup = 0;
    for (uint32_t k = 0; k < n_orig && !dup; ++k) {
        const skh_keypoint o = orig[k];
        const float r = fmaxf(o.sigma / kp.sigma, kp.sigma / o.sigma);
        dup = fm::hypotf_cr(o.x - kp.x, o.y - kp.y) <= 1.0f && r <= 1.5f;
    }
    keep[e] = dup ? 0u : 1u;
}

// The first `need` kept extras, in order: one workgroup scans the flags.
__global__ __launch_bounds__(1024) void select_kernel(const uint32_t* __restrict__ keep, uint32_t n,
                                                      uint32_t need, uint32_t* __restrict__ sel,
                                                      uint32_t* __restrict__ n_sel) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t base;
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t c0 = 0; c0 < n; c0 += 1024) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t f = i < n ? keep[i] : 0u;
        const uint64_t m = __ballot(f);
        const uint32_t pre = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        uint32_t off = base;
        for (int k = 0; k < wave; ++k) off += wsum[k];
        const uint32_t rank = off + pre;
        if (f && rank < need) sel[rank] = i;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int k = 0; k < 16; ++k) t += wsum[k];
            base += t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *n_sel = base < need ? base : need;
}

// The scan keys of the candidate records and the identity order, the radix
// sort's inputs.
__global__ __launch_bounds__(256) void cand_keys_kernel(const Cand* __restrict__ cands, uint32_t n,
                                                        uint64_t* __restrict__ keys,
                                                        uint32_t* __restrict__ idx) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    keys[i] = cands[i].key;
    idx[i] = i;
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct Buf {
    void* p = nullptr;
    size_t cap = 0;
};

// One pyramid's device state (the image, or its 2x upsample).
struct Pyramid {
    Buf base, next, g0, g1, tmp, dog, grad, cands, ncand, order_in, order, keys_in, keys,
        vals_in, vals, gset;
    GradSet hg{};
    uint32_t n_cand = 0, n_feat = 0, n_keep = 0;
};

struct Detector {
    Buf img, up, sort_tmp, keep, sel, nsel, orig_kp, res_kp, res_desc;
    Pyramid pyr[2];
};

int grow(Ctx* ctx, Buf& b, size_t bytes, const char* what) {
    if (b.cap >= bytes && b.p) return SKH_OK;
    if (b.p) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return ctx->hip(e, "hipStreamSynchronize");
        (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    want += want / 8;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) return ctx->hip(e, what);
    b.cap = want;
    return SKH_OK;
}

void free_buf(Buf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// Gaussian taps of gaussian_blur (features.rs:215-221), with glibc's expf as
// the reference computes them. Returns false if the radius exceeds kMaxR.
bool make_taps(float sigma, Taps& t) {
    float rv = std::ceil(3.0f * sigma);
    rv = rv > 1.0f ? rv : 1.0f;
    if (!(rv <= (float)kMaxR)) return false;
    const int r = (int)rv;
    t.r = r;
    for (int i = -r; i <= r; ++i) t.k[i + r] = expf(-(float)(i * i) / (2.0f * sigma * sigma));
    float s = -0.0f;
    for (int i = 0; i <= 2 * r; ++i) s = s + t.k[i];
    for (int i = 0; i <= 2 * r; ++i) t.k[i] = t.k[i] / s;
    return true;
}

// f32::powi (binary exponentiation, as LLVM lowers it).
float powi_f(float a, int b) {
    float r = 1.0f;
    for (;;) {
        if (b & 1) r = r * a;
        b /= 2;
        if (b == 0) break;
        a = a * a;
    }
    return r;
}

int launch_check(Ctx* ctx, const char* what) {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SKH_OK : ctx->hip(e, what);
}

// blur src -> dst (w x h) through tmp; with prev/dog also the DoG layer.
int blur(Ctx* ctx, const float* src, float* dst, float* tmp, int w, int h, float sigma,
         const float* prev, float* dog) {
    Taps t;
    if (!make_taps(sigma, t))
        return ctx->fail(SKH_EINVAL, "detector: Gaussian sigma %g needs more than %d taps", sigma,
                         2 * kMaxR + 1);
    const dim3 grid((w + 255) / 256, h);
    blur_h_kernel<<<grid, 256, 0, ctx->stream>>>(src, tmp, w, t);
    blur_v_kernel<<<grid, 256, 0, ctx->stream>>>(tmp, dst, w, h, t, prev, dog);
    return launch_check(ctx, "blur kernels");
}

int sort_pairs(Ctx* ctx, Detector& D, const uint64_t* k_in, uint64_t* k_out, const uint32_t* v_in,
               uint32_t* v_out, uint32_t n) {
    size_t bytes = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, k_in, k_out, v_in, v_out, n, 0, 64,
                                             ctx->stream);
    if (e != hipSuccess) return ctx->hip(e, "radix_sort_pairs (size)");
    int rc = grow(ctx, D.sort_tmp, bytes, "sort scratch");
    if (rc) return rc;
    size_t have = D.sort_tmp.cap;
    e = rocprim::radix_sort_pairs(D.sort_tmp.p, have, k_in, k_out, v_in, v_out, n, 0, 64,
                                  ctx->stream);
    return e == hipSuccess ? SKH_OK : ctx->hip(e, "radix_sort_pairs");
}

int read_u32(Ctx* ctx, const void* d, uint32_t* h) {
    int rc = ctx->hip(hipMemcpyAsync(h, d, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream),
                      "copy count");
    if (rc) return rc;
    return ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

// detect_core (features.rs:650-781) up to the sorted, truncated feature keys:
// P.n_keep features, keys/vals sorted, candidates and grad tables resident.
int detect_core(Ctx* ctx, Detector& D, Pyramid& P, const float* d_img, int w0, int h0,
                const skh_detector_cfg& cfg, float in_sigma) {
    int rc;
    const int s = (int)cfg.scales;
    P.n_cand = P.n_feat = P.n_keep = 0;
    const float kstep = powf(2.0f, 1.0f / (float)s);
    const size_t np0 = (size_t)w0 * h0;
    // octave sizes
    int nocts = 0, ws[kMaxOctaves + 1], hs[kMaxOctaves + 1];
    {
        int w = w0, h = h0;
        for (uint64_t o = 0; o < cfg.octaves; ++o) {
            if (w < 16 || h < 16) break;
            if (nocts == kMaxOctaves)
                return ctx->fail(SKH_EINVAL, "detector: more than %d octaves", kMaxOctaves);
            ws[nocts] = w;
            hs[nocts] = h;
            ++nocts;
            w /= 2;
            h /= 2;
        }
    }
    size_t grad_px = 0, cand_cap = 64;
    for (int o = 0; o < nocts; ++o) {
        grad_px += (size_t)s * ws[o] * hs[o];
        cand_cap += (size_t)s * (ws[o] - 2) * (hs[o] - 2) / 2;
    }
    if (cand_cap >= (1u << 27))
        return ctx->fail(SKH_EINVAL, "detector: image too large (%zu candidate slots)", cand_cap);
    if ((rc = grow(ctx, P.base, sizeof(float) * np0, "detector base")) ||
        (rc = grow(ctx, P.next, sizeof(float) * np0, "detector base")) ||
        (rc = grow(ctx, P.g0, sizeof(float) * np0, "detector gauss")) ||
        (rc = grow(ctx, P.g1, sizeof(float) * np0, "detector gauss")) ||
        (rc = grow(ctx, P.tmp, sizeof(float) * np0, "detector tmp")) ||
        (rc = grow(ctx, P.dog, sizeof(float) * np0 * (s + 2), "detector dog")) ||
        (rc = grow(ctx, P.grad, sizeof(float2) * std::max<size_t>(grad_px, 1), "detector grad")) ||
        (rc = grow(ctx, P.cands, sizeof(Cand) * cand_cap, "detector candidates")) ||
        (rc = grow(ctx, P.ncand, 2 * sizeof(uint32_t), "detector counts")) ||
        (rc = grow(ctx, P.gset, sizeof(GradSet), "detector grad set")))
        return rc;
    zero_u32_kernel<<<1, 64, 0, ctx->stream>>>((uint32_t*)P.ncand.p, 2);
    if ((rc = launch_check(ctx, "zero counts"))) return rc;

    float sig_init = cfg.sigma0 * cfg.sigma0 - in_sigma * in_sigma;
    sig_init = sig_init > 0.01f ? sig_init : 0.01f;
    float* base = (float*)P.base.p;
    float* next = (float*)P.next.p;
    if (nocts > 0 &&
        (rc = blur(ctx, d_img, base, (float*)P.tmp.p, w0, h0, std::sqrt(sig_init), nullptr,
                   nullptr)))
        return rc;
    const float edge_thr = powi_f(cfg.edge_ratio + 1.0f, 2) / cfg.edge_ratio;
    memset(&P.hg, 0, sizeof P.hg);
    float2* gp = (float2*)P.grad.p;
    for (int o = 0; o < nocts; ++o) {
        const int w = ws[o], h = hs[o];
        const size_t np = (size_t)w * h;
        const dim3 grid((w + 255) / 256, h);
        P.hg.w[o] = w;
        P.hg.h[o] = h;
        DogSet ds{};
        float* dog = (float*)P.dog.p;
        for (int i = 0; i < s + 2; ++i) ds.d[i] = dog + (size_t)i * np;
        // gauss[0] = base; gauss[i] = blur(gauss[i-1]); dog[i-1] = gauss[i] - gauss[i-1]
        const float* prev = base;
        float* gb[2] = {(float*)P.g0.p, (float*)P.g1.p};
        for (int i = 1; i < s + 3; ++i) {
            const float sp = cfg.sigma0 * powi_f(kstep, i - 1);
            const float cur = sp * kstep;
            float* dst = gb[i & 1];
            if ((rc = blur(ctx, prev, dst, (float*)P.tmp.p, w, h, std::sqrt(cur * cur - sp * sp),
                           prev, dog + (size_t)(i - 1) * np)))
                return rc;
            if (i <= s) {
                P.hg.g[o][i] = gp;
                grad_kernel<<<grid, 256, 0, ctx->stream>>>(dst, gp, w, h);
                gp += np;
            }
            if (i == s && o + 1 < nocts) {
                const dim3 g2((w / 2 + 255) / 256, h / 2);
                downsample_kernel<<<g2, 256, 0, ctx->stream>>>(dst, w, next, w / 2, h / 2);
            }
            prev = dst;
        }
        const dim3 eg((w - 2 + 255) / 256, h - 2, s);
        extrema_kernel<<<eg, 256, 0, ctx->stream>>>(ds, w, h, s, o, cfg.contrast, edge_thr,
                                                    (Cand*)P.cands.p, (uint32_t*)P.ncand.p,
                                                    (uint32_t)cand_cap);
        if ((rc = launch_check(ctx, "pyramid kernels"))) return rc;
        std::swap(base, next);
    }
    if ((rc = read_u32(ctx, P.ncand.p, &P.n_cand))) return rc;
    if (P.n_cand > cand_cap)
        return ctx->fail(SKH_EHIP, "detector: %u candidates exceed the bound %zu", P.n_cand,
                         cand_cap);
    if (P.n_cand == 0) return SKH_OK;
    const uint32_t nc = P.n_cand;
    // candidates into emission order: sort their scan keys
    if ((rc = grow(ctx, P.keys_in, sizeof(uint64_t) * nc * kMaxAngles, "detector keys")) ||
        (rc = grow(ctx, P.keys, sizeof(uint64_t) * nc * kMaxAngles, "detector keys")) ||
        (rc = grow(ctx, P.vals_in, sizeof(uint32_t) * nc * kMaxAngles, "detector vals")) ||
        (rc = grow(ctx, P.vals, sizeof(uint32_t) * nc * kMaxAngles, "detector vals")) ||
        (rc = grow(ctx, P.order_in, sizeof(uint32_t) * nc, "detector order")) ||
        (rc = grow(ctx, P.order, sizeof(uint32_t) * nc, "detector order")))
        return rc;
    cand_keys_kernel<<<(nc + 255) / 256, 256, 0, ctx->stream>>>(
        (const Cand*)P.cands.p, nc, (uint64_t*)P.keys_in.p, (uint32_t*)P.order_in.p);
    if ((rc = launch_check(ctx, "cand_keys_kernel"))) return rc;
    if ((rc = sort_pairs(ctx, D, (const uint64_t*)P.keys_in.p, (uint64_t*)P.keys.p,
                         (const uint32_t*)P.order_in.p, (uint32_t*)P.order.p, nc)))
        return rc;
    if ((rc = ctx->hip(hipMemcpyAsync(P.gset.p, &P.hg, sizeof(GradSet), hipMemcpyHostToDevice,
                                      ctx->stream),
                       "copy grad set")))
        return rc;
    orient_kernel<<<nc, 64, 0, ctx->stream>>>(
        (const GradSet*)P.gset.p, (Cand*)P.cands.p, (const uint32_t*)P.order.p, nc, cfg.sigma0,
        kstep, (uint64_t*)P.keys_in.p, (uint32_t*)P.vals_in.p, (uint32_t*)P.ncand.p + 1);
    if ((rc = launch_check(ctx, "orient_kernel"))) return rc;
    if ((rc = read_u32(ctx, (uint32_t*)P.ncand.p + 1, &P.n_feat))) return rc;
    if (P.n_feat == 0) return SKH_OK;
    if ((rc = sort_pairs(ctx, D, (const uint64_t*)P.keys_in.p, (uint64_t*)P.keys.p,
                         (const uint32_t*)P.vals_in.p, (uint32_t*)P.vals.p, P.n_feat)))
        return rc;
    P.n_keep = (uint32_t)std::min<uint64_t>(P.n_feat, cfg.max_features);
    return SKH_OK;
}

DescJob job_of(const Pyramid& P, uint32_t n, uint32_t out0, const uint32_t* sel, int upscaled) {
    DescJob J;
    J.keys = (const uint64_t*)P.keys.p;
    J.angles = (const uint32_t*)P.vals.p;
    J.order = (const uint32_t*)P.order.p;
    J.cands = (const Cand*)P.cands.p;
    J.G = (const GradSet*)P.gset.p;
    J.sel = sel;
    J.n = n;
    J.out0 = out0;
    J.upscaled = upscaled;
    return J;
}

Detector& detector_of(Ctx* ctx) {
    if (!ctx->detector) ctx->detector = new Detector();
    return *static_cast<Detector*>(ctx->detector);
}

// detect_and_describe (features.rs:594-648) into device outputs d_kps (cap
// keypoints) and d_desc (cap x 128, may be NULL: keypoints only).
int detect_dev(Ctx* ctx, const float* d_img, size_t w, size_t h, const skh_detector_cfg& cfg,
               skh_keypoint* d_kps, float* d_desc, size_t cap, size_t* out_len) {
    Detector& D = detector_of(ctx);
    int rc;
    *out_len = 0;
    if (cfg.scales == 0 || cfg.scales + 2 > (uint64_t)kMaxLayers)
        return ctx->fail(SKH_EINVAL, "detector: scales = %llu outside 1..%d",
                         (unsigned long long)cfg.scales, kMaxLayers - 2);
    if (w == 0 || h == 0) return SKH_OK;
    // the upsampled image (2w x 2h) must fit the grids: rows on gridDim.y
    if (w > (1u << 22) || h > 32767)
        return ctx->fail(SKH_EINVAL, "detector: image %zu x %zu too large", w, h);
    Pyramid& P = D.pyr[0];
    if ((rc = detect_core(ctx, D, P, d_img, (int)w, (int)h, cfg, 0.5f))) return rc;
    const uint32_t n0 = P.n_keep;
    const size_t longer = std::max(w, h);
    const bool upscale = !(cfg.upscale_below == 0 || longer >= cfg.upscale_below ||
                           n0 >= cfg.max_features ||
                           (!cfg.upscale_fill && (uint64_t)n0 * 2 >= cfg.max_features));
    uint32_t n_add = 0;
    Pyramid& U = D.pyr[1];
    if (upscale) {
        // the originals' keypoints are needed by the duplicate test
        if (n0 > 0) {
            if ((rc = grow(ctx, D.orig_kp, sizeof(skh_keypoint) * n0, "detector keypoints")))
                return rc;
            const DescJob J = job_of(P, n0, 0, nullptr, 0);
            describe_kernel<<<n0, 64, 0, ctx->stream>>>(J, (skh_keypoint*)D.orig_kp.p, nullptr);
            if ((rc = launch_check(ctx, "describe_kernel (keypoints)"))) return rc;
        }
        if ((rc = grow(ctx, D.up, sizeof(float) * 4 * w * h, "detector upsample"))) return rc;
        const dim3 ug((2 * (int)w + 255) / 256, 2 * (int)h);
        upsample_kernel<<<ug, 256, 0, ctx->stream>>>(d_img, (int)w, (int)h, (float*)D.up.p);
        if ((rc = launch_check(ctx, "upsample_kernel"))) return rc;
        if ((rc = detect_core(ctx, D, U, (const float*)D.up.p, 2 * (int)w, 2 * (int)h, cfg, 1.0f)))
            return rc;
        const uint32_t need = (uint32_t)(cfg.max_features - n0);
        if (U.n_keep > 0 && need > 0) {
            if ((rc = grow(ctx, D.keep, sizeof(uint32_t) * U.n_keep, "detector keep")) ||
                (rc = grow(ctx, D.sel, sizeof(uint32_t) * U.n_keep, "detector sel")) ||
                (rc = grow(ctx, D.nsel, sizeof(uint32_t), "detector nsel")))
                return rc;
            const DescJob J = job_of(U, U.n_keep, 0, nullptr, 1);
            dedup_kernel<<<(U.n_keep + 255) / 256, 256, 0, ctx->stream>>>(
                J, (const skh_keypoint*)D.orig_kp.p, n0, (uint32_t*)D.keep.p);
            select_kernel<<<1, 1024, 0, ctx->stream>>>((const uint32_t*)D.keep.p, U.n_keep, need,
                                                       (uint32_t*)D.sel.p, (uint32_t*)D.nsel.p);
            if ((rc = launch_check(ctx, "dedup / select"))) return rc;
            if ((rc = read_u32(ctx, D.nsel.p, &n_add))) return rc;
        }
    }
    const size_t total = (size_t)n0 + n_add;
    *out_len = total;
    if (total > cap)
        return ctx->fail(SKH_EINVAL, "detector: %zu features exceed the capacity %zu", total, cap);
    if (n0 > 0) {
        const DescJob J = job_of(P, n0, 0, nullptr, 0);
        describe_kernel<<<n0, 64, 0, ctx->stream>>>(J, d_kps, d_desc);
    }
    if (n_add > 0) {
        const DescJob J = job_of(U, n_add, n0, (const uint32_t*)D.sel.p, 1);
        describe_kernel<<<n_add, 64, 0, ctx->stream>>>(J, d_kps, d_desc);
    }
    return launch_check(ctx, "describe_kernel");
}

int check_cfg(Ctx* ctx, const skh_detector_cfg* cfg) {
    if (!cfg) return ctx->fail(SKH_EINVAL, "detector: NULL config");
    return SKH_OK;
}

}  // namespace

void destroy_detector(void* p) {
    if (!p) return;
    Detector* D = static_cast<Detector*>(p);
    Buf* bufs[] = {&D->img,  &D->up,      &D->sort_tmp, &D->keep,    &D->sel,
                   &D->nsel, &D->orig_kp, &D->res_kp,   &D->res_desc};
    for (Buf* b : bufs) free_buf(*b);
    for (Pyramid& P : D->pyr) {
        Buf* pb[] = {&P.base,  &P.next, &P.g0,      &P.g1,    &P.tmp,     &P.dog,
                     &P.grad,  &P.cands, &P.ncand,  &P.order_in, &P.order, &P.keys_in,
                     &P.keys,  &P.vals_in, &P.vals, &P.gset};
        for (Buf* b : pb) free_buf(*b);
    }
    delete D;
}

}  // namespace skh

using namespace skh;

extern "C" void skh_detector_cfg_default(skh_detector_cfg* c) {
    if (!c) return;
    c->octaves = 4;
    c->scales = 3;
    c->sigma0 = 1.6f;
    c->contrast = 0.01f;
    c->edge_ratio = 10.0f;
    c->upscale_fill = 0;
    c->max_features = 8192;
    c->upscale_below = 800;
}

extern "C" int skh_detect_and_describe_dev(skh_ctx* ctx, const float* d_img, size_t w, size_t h,
                                           const skh_detector_cfg* cfg, skh_keypoint* d_kps,
                                           float* d_desc, size_t cap, size_t* out_len) {
    if (!ctx || !out_len || (!d_img && w * h > 0) || (cap > 0 && !d_kps))
        return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    int rc = check_cfg(ctx, cfg);
    if (rc) return rc;
    rc = ctx->hip(hipSetDevice(ctx->device), "hipSetDevice");
    if (rc) return rc;
    return detect_dev(ctx, d_img, w, h, *cfg, d_kps, d_desc, cap, out_len);
}

extern "C" int skh_detect_and_describe(skh_ctx* ctx, const float* img, size_t w, size_t h,
                                       const skh_detector_cfg* cfg, skh_keypoint* kps, float* desc,
                                       size_t cap, size_t* out_len) {
    if (!ctx || !out_len || (!img && w * h > 0) || (cap > 0 && !kps)) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    int rc = check_cfg(ctx, cfg);
    if (rc) return rc;
    if ((rc = ctx->hip(hipSetDevice(ctx->device), "hipSetDevice"))) return rc;
    Detector& D = detector_of(ctx);
    const size_t np = w * h;
    if ((rc = grow(ctx, D.img, sizeof(float) * std::max<size_t>(np, 1), "detector image")))
        return rc;
    if (np > 0 && (rc = ctx->hip(hipMemcpyAsync(D.img.p, img, sizeof(float) * np,
                                                hipMemcpyHostToDevice, ctx->stream),
                                 "copy image")))
        return rc;
    // outputs land in context buffers, then on the host
    static_assert(sizeof(skh_keypoint) == 20, "skh_keypoint is five f32");
    const size_t dcap = std::max<size_t>(std::min<size_t>(cap, cfg->max_features), 1);
    if ((rc = grow(ctx, D.res_desc, sizeof(float) * 128 * dcap, "detector descriptors")) ||
        (rc = grow(ctx, D.res_kp, sizeof(skh_keypoint) * dcap, "detector keypoints")))
        return rc;
    rc = detect_dev(ctx, (const float*)D.img.p, w, h, *cfg, (skh_keypoint*)D.res_kp.p,
                    desc ? (float*)D.res_desc.p : nullptr, dcap, out_len);
    if (rc) return rc;
    const size_t n = *out_len;
    if (n > cap) return ctx->fail(SKH_EINVAL, "detector: %zu features exceed the capacity %zu", n, cap);
    if (n > 0) {
        if ((rc = ctx->hip(hipMemcpyAsync(kps, D.res_kp.p, sizeof(skh_keypoint) * n,
                                          hipMemcpyDeviceToHost, ctx->stream),
                           "copy keypoints")))
            return rc;
        if (desc && (rc = ctx->hip(hipMemcpyAsync(desc, D.res_desc.p, sizeof(float) * 128 * n,
                                                  hipMemcpyDeviceToHost, ctx->stream),
                                   "copy descriptors")))
            return rc;
    }
    return ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}
