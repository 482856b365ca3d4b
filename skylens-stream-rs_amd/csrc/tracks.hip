// build_tracks (crates/core/src/tracks.rs:190-508): multi-view tracks from
// verified pairwise matches.
//
// The per-match work, the displacement filter of the Split policy
// (tracks.rs:519-670: 24 nearest matches on a grid, three robust affine
// fits of the local displacement field, a median of 24 residuals), runs on
// the device, one lane per match of every pair at once. The graph stages
// that follow (edge votes, the image-aware union-find, the component rounds,
// the track cap) are sequential merges over sorted lists and stay on the
// host. Both follow the reference step for step; the 24 nearest candidates
// are taken in (squared distance, index) order where the reference's
// select_nth_unstable leaves the order unspecified.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "skh_ctx.hpp"
#include "skh_internal.hpp"
#include "skylens_hip.h"

namespace skh {
namespace {

constexpr double kDispTolerance = 40.0;  // tracks.rs:6-11
constexpr double kDispFloor = 0.25;
constexpr double kDispCap = 3.0;
constexpr double kRefW = 960.0, kRefH = 540.0;
constexpr int kNeighbors = 24;
constexpr int kVoteRounds = 3;  // tracks.rs:158

// One pair's matches for the displacement filter: points [p0, p0 + n) of the
// flat arrays, its grid cell and tolerance.
struct DispPair {
    uint32_t p0, n;
    double cell, tol;
};
// Grid key of a point and its index within the pair; sorted per pair by
// (kx, ky, i).
struct CellKey {
    int64_t kx, ky;
    uint32_t i, pad;
};

// `as i64` of a float: saturating, NaN -> 0.
__host__ __device__ inline int64_t sat_i64(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}

// (d2, j) before (e2, k): squared distance, then index.
__device__ inline bool cand_less(double d2, uint32_t j, double e2, uint32_t k) {
    return d2 < e2 || (d2 == e2 && j < k);
}

// fit_affine (tracks.rs:672-689): least squares of dx = a0 + a1 u + a2 v,
// dy = b0 + b1 u + b2 v over the kept neighbours, through the 3x3 normal
// matrix's determinant and adjugate (nalgebra's 3x3 forms).
__device__ bool fit_affine(const double (*nb)[4], uint32_t keep, int n, double* f) {
    double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, rx[3] = {0, 0, 0}, ry[3] = {0, 0, 0};
    for (int k = 0; k < n; ++k) {
        if (!((keep >> k) & 1u)) continue;
        const double v[3] = {1.0, nb[k][0], nb[k][1]};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
#pragma unroll
            for (int c = 0; c < 3; ++c) m[r][c] = m[r][c] + v[r] * v[c];
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) rx[r] = rx[r] + v[r] * nb[k][2];
#pragma unroll
        for (int r = 0; r < 3; ++r) ry[r] = ry[r] + v[r] * nb[k][3];
    }
    const double m11 = m[0][0], m12 = m[0][1], m13 = m[0][2];
    const double m21 = m[1][0], m22 = m[1][1], m23 = m[1][2];
    const double m31 = m[2][0], m32 = m[2][1], m33 = m[2][2];
    const double c1 = m22 * m33 - m32 * m23;
    const double c2 = m21 * m33 - m31 * m23;
    const double c3 = m21 * m32 - m31 * m22;
    const double det = m11 * c1 - m12 * c2 + m13 * c3;
    if (det != det || det <= 1e-4 * (m11 * m11 * m11) || det == 0.0) return false;
    double inv[3][3];
    inv[0][0] = c1 / det;
    inv[0][1] = (m13 * m32 - m33 * m12) / det;
    inv[0][2] = (m12 * m23 - m22 * m13) / det;
    inv[1][0] = -c2 / det;
    inv[1][1] = (m11 * m33 - m31 * m13) / det;
    inv[1][2] = (m13 * m21 - m23 * m11) / det;
    inv[2][0] = c3 / det;
    inv[2][1] = (m12 * m31 - m32 * m11) / det;
    inv[2][2] = (m11 * m22 - m21 * m12) / det;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        f[r] = inv[r][0] * rx[0] + inv[r][1] * rx[1] + inv[r][2] * rx[2];
        f[3 + r] = inv[r][0] * ry[0] + inv[r][1] * ry[1] + inv[r][2] * ry[2];
    }
    return true;
}

// The displacement filter for one match (a lane): pts[x] = (x_a, y_a,
// x_b - x_a, y_b - y_a).
__global__ __launch_bounds__(64) void displacement_kernel(const DispPair* __restrict__ pairs,
                                                          const uint32_t* __restrict__ point_pair,
                                                          const double4* __restrict__ pts,
                                                          const CellKey* __restrict__ keys,
                                                          uint32_t n_points, double floor_,
                                                          double cap_, uint8_t* __restrict__ flags) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n_points) return;
    const DispPair P = pairs[point_pair[x]];
    const uint32_t i = x - P.p0;
    const double4 q = pts[x];
    const int64_t cx = sat_i64(floor(q.x / P.cell)), cy = sat_i64(floor(q.y / P.cell));
    const CellKey* K = keys + P.p0;
    // entries of grid cell (gx, gy): [first >= (gx, gy), first > (gx, gy))
    auto cell_start = [&](int64_t gx, int64_t gy) {
        uint32_t lo = 0, hi = P.n;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            const bool less = K[mid].kx < gx || (K[mid].kx == gx && K[mid].ky < gy);
            if (less) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    // rings 1..3 until one holds at least 24 candidates (tracks.rs:603-621)
    int ring = 1;
    uint32_t count = 0;
    for (; ring <= 3; ++ring) {
        count = 0;
        for (int64_t gx = cx - ring; gx <= cx + ring; ++gx)
            for (int64_t gy = cy - ring; gy <= cy + ring; ++gy)
                for (uint32_t e = cell_start(gx, gy); e < P.n && K[e].kx == gx && K[e].ky == gy; ++e)
                    count += K[e].i != i ? 1u : 0u;
        if (count >= (uint32_t)kNeighbors) break;
    }
    if (ring > 3) ring = 3;
    uint8_t flag = 0;
    if (count >= 8) {
        // the candidates: in scan order when at most 24, else the 24 smallest by
        // (squared distance, index), kept sorted by insertion
        double cd[kNeighbors];
        uint32_t cj[kNeighbors];
        int nc = 0;
        const bool select = count > (uint32_t)kNeighbors;
        for (int64_t gx = cx - ring; gx <= cx + ring; ++gx)
            for (int64_t gy = cy - ring; gy <= cy + ring; ++gy)
                for (uint32_t e = cell_start(gx, gy); e < P.n && K[e].kx == gx && K[e].ky == gy; ++e) {
                    const uint32_t j = K[e].i;
                    if (j == i) continue;
                    const double4 o = pts[P.p0 + j];
                    const double ex = o.x - q.x, ey = o.y - q.y;
                    const double d2 = ex * ex + ey * ey;
                    if (!select) {
                        cd[nc] = d2;
                        cj[nc] = j;
                        ++nc;
                        continue;
                    }
                    if (nc == kNeighbors && !cand_less(d2, j, cd[nc - 1], cj[nc - 1])) continue;
                    int s = nc < kNeighbors ? nc++ : kNeighbors - 1;
                    while (s > 0 && cand_less(d2, j, cd[s - 1], cj[s - 1])) {
                        cd[s] = cd[s - 1];
                        cj[s] = cj[s - 1];
                        --s;
                    }
                    cd[s] = d2;
                    cj[s] = j;
                }
        double nb[kNeighbors][4];
        for (int k = 0; k < nc; ++k) {
            const double4 o = pts[P.p0 + cj[k]];
            nb[k][0] = (o.x - q.x) / P.cell;
            nb[k][1] = (o.y - q.y) / P.cell;
            nb[k][2] = o.z - q.z;
            nb[k][3] = o.w - q.w;
        }
        uint32_t keep = nc >= 32 ? 0xffffffffu : ((1u << nc) - 1u);
        double f[6] = {0, 0, 0, 0, 0, 0}, s = 0.0;
        bool have = false;
        for (int round = 0; round < 3; ++round) {
            double g[6];
            if (!fit_affine(nb, keep, nc, g)) break;
            double res[kNeighbors], srt[kNeighbors];
            for (int k = 0; k < nc; ++k) {
                const double rx = nb[k][2] - (g[0] + g[1] * nb[k][0] + g[2] * nb[k][1]);
                const double ry = nb[k][3] - (g[3] + g[4] * nb[k][0] + g[5] * nb[k][1]);
                res[k] = sqrt(rx * rx + ry * ry);
                // insertion into ascending order (non-negative residuals)
                int t = k;
                while (t > 0 && srt[t - 1] > res[k]) {
                    srt[t] = srt[t - 1];
                    --t;
                }
                srt[t] = res[k];
            }
            s = 1.4826 * srt[nc / 2];
#pragma unroll
            for (int k = 0; k < 6; ++k) f[k] = g[k];
            have = true;
            if (round == 2) break;
            const double cut = fmax(3.0 * s, P.tol * 0.25);
            keep = 0;
            int kept = 0;
            for (int k = 0; k < nc; ++k)
                if (res[k] <= cut) {
                    keep |= 1u << k;
                    ++kept;
                }
            if (kept < 6) break;
        }
        if (have) {
            const double r = sqrt(f[0] * f[0] + f[3] * f[3]);
            double lim = 4.0 * s;
            const double lo_l = P.tol * floor_, hi_l = cap_ * P.tol;
            if (lim < lo_l) lim = lo_l;
            if (lim > hi_l) lim = hi_l;
            flag = r > lim ? 1 : 0;
        }
    }
    flags[x] = flag;
}

// ------------------------------------------------------------ host: graph

// f(i) for every i in [0, n), on up to 16 host threads in contiguous blocks
// (f must only write state of its own i).
template <class F>
void parallel_for(size_t n, F&& f) {
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t T = std::min<size_t>(std::min<size_t>(hw, 16), (n + 255) / 256);
    if (T <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T);
    for (size_t t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            for (size_t i = n * t / T, e = n * (t + 1) / T; i < e; ++i) f(i);
        });
    for (auto& x : th) x.join();
}

struct UnionFind {  // tracks.rs:160-188
    std::vector<size_t> parent;
    explicit UnionFind(size_t n) : parent(n) {
        for (size_t x = 0; x < n; ++x) parent[x] = x;
    }
    size_t find(size_t x) {
        size_t r = x;
        while (parent[r] != r) r = parent[r];
        while (parent[x] != r) {
            const size_t next = parent[x];
            parent[x] = r;
            x = next;
        }
        return r;
    }
    size_t link(size_t ra, size_t rb) {
        const size_t lo = std::min(ra, rb), hi = std::max(ra, rb);
        parent[hi] = lo;
        return lo;
    }
};

// Common elements of two ascending lists (tracks.rs:750-764).
size_t common_count(const size_t* a, size_t na, const size_t* b, size_t nb) {
    size_t i = 0, j = 0, c = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) ++i;
        else if (b[j] < a[i]) ++j;
        else {
            ++c;
            ++i;
            ++j;
        }
    }
    return c;
}

// The sorted image lists of the union-find's roots, as regions of one arena:
// root r's list is arena[off[r], off[r] + len[r]) inside a region of cap[r]
// entries; len[r] = 0 for a lone node, whose list is its own image.
struct ImageLists {
    std::vector<size_t> off, len, cap;
    std::vector<uint32_t> arena;  // image ids (below 2^32, checked at entry)
    explicit ImageLists(size_t n) : off(n, 0), len(n, 0), cap(n, 0) {}
};

// Joins the components of u and v if their image sets are disjoint
// (tracks.rs:692-724).
bool try_join(UnionFind& uf, ImageLists& L, const std::vector<size_t>& node_image, size_t u,
              size_t v) {
    const size_t ru = uf.find(u), rv = uf.find(v);
    if (ru == rv) return true;
    const uint32_t su = (uint32_t)node_image[ru], sv = (uint32_t)node_image[rv];
    const uint32_t* iu = L.len[ru] ? L.arena.data() + L.off[ru] : &su;
    const uint32_t* iv = L.len[rv] ? L.arena.data() + L.off[rv] : &sv;
    const size_t nu = L.len[ru] ? L.len[ru] : 1, nv = L.len[rv] ? L.len[rv] : 1;
    // any common image refuses the join (tracks.rs:715 counts them; only zero
    // or not is used)
    for (size_t i = 0, j = 0; i < nu && j < nv;) {
        if (iu[i] < iv[j]) ++i;
        else if (iv[j] < iu[i]) ++j;
        else return false;
    }
    // merged into a reused buffer, then into the larger of the two regions when
    // it holds the result, else into a new region (twice the size) at the end
    // of the arena; the new root keeps it
    static thread_local std::vector<uint32_t> merged;
    merged.resize(nu + nv);
    std::merge(iu, iu + nu, iv, iv + nv, merged.begin());
    const size_t r = uf.link(ru, rv), o = r == ru ? rv : ru;
    const size_t big = L.cap[o] > L.cap[r] ? o : r;
    size_t off = L.off[big], cap = L.cap[big];
    if (cap < nu + nv) {
        off = L.arena.size();
        cap = std::max<size_t>(2 * (nu + nv), 4);
        L.arena.resize(L.arena.size() + cap);
    }
    std::copy(merged.begin(), merged.end(), L.arena.begin() + off);
    L.off[r] = off;
    L.cap[r] = cap;
    L.len[r] = nu + nv;
    L.len[o] = L.cap[o] = 0;
    return true;
}

// Stable sort of v by key(x) descending (a counting sort: the keys are counts
// bounded by the input size).
template <class T, class K>
void stable_desc(std::vector<T>& v, K key) {
    size_t top = 0;
    for (const T& x : v) top = std::max(top, (size_t)key(x));
    std::vector<size_t> at(top + 2, 0);  // at[top - k]: first slot of key k
    for (const T& x : v) at[top - key(x) + 1]++;
    for (size_t k = 0; k <= top; ++k) at[k + 1] += at[k];
    std::vector<T> out(v.size());
    for (const T& x : v) out[at[top - key(x)]++] = x;
    v.swap(out);
}

// Ascending order of (first, second) pairs whose first is below n: a counting
// sort on first, then each first's run by second.
void sort_pairs(std::vector<std::pair<size_t, size_t>>& v, size_t n) {
    std::vector<size_t> at(n + 1, 0);
    for (const auto& e : v) at[e.first + 1]++;
    for (size_t x = 0; x < n; ++x) at[x + 1] += at[x];
    std::vector<std::pair<size_t, size_t>> out(v.size());
    {
        std::vector<size_t> fill(at.begin(), at.end() - 1);
        for (const auto& e : v) out[fill[e.first]++] = e;
    }
    for (size_t x = 0; x < n; ++x)
        if (at[x + 1] - at[x] > 1) std::sort(out.begin() + at[x], out.begin() + at[x + 1]);
    v.swap(out);
}

struct CompPair {
    size_t a, b, count, sup;
};

// Pairs of distinct components joined by zero-support edges, with their edge
// counts, ascending (tracks.rs:727-747).
std::vector<CompPair> component_pairs(const std::vector<std::array<size_t, 3>>& zero,
                                      UnionFind& uf) {
    std::vector<std::pair<size_t, size_t>> groups;
    for (const auto& e : zero) {
        const size_t ru = uf.find(e[1]), rv = uf.find(e[2]);
        if (ru != rv) groups.emplace_back(std::min(ru, rv), std::max(ru, rv));
    }
    sort_pairs(groups, uf.parent.size());
    std::vector<CompPair> out;
    for (const auto& g : groups) {
        if (!out.empty() && out.back().a == g.first && out.back().b == g.second) out.back().count++;
        else out.push_back({g.first, g.second, 1, 0});
    }
    return out;
}

using Edge = std::pair<size_t, size_t>;

// The Split policy's votes and joins (tracks.rs:259-465) over the sorted,
// deduplicated edges; edges keeps the uncut ones.
void split_policy(std::vector<Edge>& edges, size_t n, const std::vector<size_t>& node_image,
                  UnionFind& uf, skh_track_stats& st) {
    // adjacency CSR; each node's neighbours ascending (edges are sorted by (u, v))
    std::vector<size_t> start(n + 1, 0);
    for (const Edge& e : edges) {
        start[e.first + 1]++;
        start[e.second + 1]++;
    }
    for (size_t i = 0; i < n; ++i) start[i + 1] += start[i];
    std::vector<size_t> flat(2 * edges.size()), fe(2 * edges.size());
    {
        std::vector<size_t> fill(start.begin(), start.end() - 1);
        for (size_t e = 0; e < edges.size(); ++e) {
            const auto [u, v] = edges[e];
            flat[fill[u]] = v;
            fe[fill[u]++] = e;
            flat[fill[v]] = u;
            fe[fill[v]++] = e;
        }
    }
    auto support = [&](size_t u, size_t v) {
        return common_count(flat.data() + start[u], start[u + 1] - start[u], flat.data() + start[v],
                            start[v + 1] - start[v]);
    };
    for (int round = 0; round < kVoteRounds; ++round) {
        std::vector<size_t> sup(edges.size());
        parallel_for(edges.size(), [&](size_t e) { sup[e] = support(edges[e].first, edges[e].second); });
        std::vector<char> cut(edges.size(), 0);
        size_t cuts = 0;
        for (size_t x = 0; x < n; ++x) {
            // runs of neighbours in the same image (node ids are contiguous per image)
            for (size_t i = start[x], hi = start[x + 1]; i < hi;) {
                const size_t img = node_image[flat[i]];
                size_t j = i + 1;
                while (j < hi && node_image[flat[j]] == img) ++j;
                if (j - i > 1) {
                    size_t best = 0, top = 0;
                    for (size_t k = i; k < j; ++k) best = std::max(best, sup[fe[k]]);
                    for (size_t k = i; k < j; ++k) top += sup[fe[k]] == best;
                    for (size_t k = i; k < j; ++k) {
                        const size_t e = fe[k];
                        const bool drop = sup[e] < best ? best > 0 : (top > 1 && best > 0);
                        if (drop && !cut[e]) {
                            cut[e] = 1;
                            ++cuts;
                        }
                    }
                }
                i = j;
            }
        }
        if (cuts == 0) break;
        st.conflicts += cuts;
        st.vote_cuts += cuts;
        std::vector<size_t> newid(edges.size(), SIZE_MAX);
        size_t k = 0;
        for (size_t e = 0; e < edges.size(); ++e)
            if (!cut[e]) newid[e] = k++;
        std::vector<Edge> kept;
        kept.reserve(k);
        for (size_t e = 0; e < edges.size(); ++e)
            if (!cut[e]) kept.push_back(edges[e]);
        edges.swap(kept);
        size_t w = 0, r = 0;
        for (size_t x = 0; x < n; ++x) {
            const size_t end = start[x + 1];
            start[x] = w;
            for (; r < end; ++r)
                if (!cut[fe[r]]) {
                    flat[w] = flat[r];
                    fe[w] = newid[fe[r]];
                    ++w;
                }
        }
        start[n] = w;
        flat.resize(w);
        fe.resize(w);
    }
    // (support, u, v): support descending, then (u, v) ascending (the edges are
    // in ascending (u, v) order and the sort on the support is stable)
    std::vector<std::array<size_t, 3>> order(edges.size());
    parallel_for(edges.size(), [&](size_t e) {
        order[e] = {support(edges[e].first, edges[e].second), edges[e].first, edges[e].second};
    });
    stable_desc(order, [](const std::array<size_t, 3>& x) { return x[0]; });
    ImageLists images(n);
    size_t split = order.size();
    for (size_t e = 0; e < order.size(); ++e)
        if (order[e][0] == 0) {
            split = e;
            break;
        }
    for (size_t e = 0; e < split; ++e)
        if (!try_join(uf, images, node_image, order[e][1], order[e][2])) st.conflicts++;
    const std::vector<std::array<size_t, 3>> zero(order.begin() + split, order.end());
    // dense[r]: the round's index of component r (SIZE_MAX when absent), so that
    // a round's work is proportional to its component pairs, not to n
    std::vector<size_t> dense(n, SIZE_MAX), ids;
    for (int round = 0; round <= 8; ++round) {
        std::vector<CompPair> counted = component_pairs(zero, uf);
        if (counted.empty()) break;
        // component graph CSR, each list sorted; support = common neighbour
        // components (the lists hold dense ids: any one order of both lists
        // gives the same count)
        ids.clear();
        for (const CompPair& c : counted)
            for (const size_t r : {c.a, c.b})
                if (dense[r] == SIZE_MAX) {
                    dense[r] = ids.size();
                    ids.push_back(r);
                }
        const size_t m = ids.size();
        std::vector<size_t> cstart(m + 1, 0);
        for (const CompPair& c : counted) {
            cstart[dense[c.a] + 1]++;
            cstart[dense[c.b] + 1]++;
        }
        for (size_t i = 0; i < m; ++i) cstart[i + 1] += cstart[i];
        std::vector<size_t> cfill(cstart.begin(), cstart.end() - 1), cflat(2 * counted.size());
        for (const CompPair& c : counted) {
            cflat[cfill[dense[c.a]]++] = dense[c.b];
            cflat[cfill[dense[c.b]]++] = dense[c.a];
        }
        for (size_t i = 0; i < m; ++i)
            if (cstart[i + 1] - cstart[i] > 1)
                std::sort(cflat.begin() + cstart[i], cflat.begin() + cstart[i + 1]);
        for (CompPair& c : counted) {
            const size_t a = dense[c.a], b = dense[c.b];
            c.sup = common_count(cflat.data() + cstart[a], cstart[a + 1] - cstart[a],
                                 cflat.data() + cstart[b], cstart[b + 1] - cstart[b]);
        }
        for (const size_t r : ids) dense[r] = SIZE_MAX;
        // count descending, then support descending, then (a, b) ascending
        // (component_pairs lists them by (a, b); both sorts are stable)
        stable_desc(counted, [](const CompPair& x) { return x.sup; });
        stable_desc(counted, [](const CompPair& x) { return x.count; });
        const bool last = round == 8;
        size_t merged = 0;
        for (const CompPair& c : counted) {
            const bool strong = c.count >= 2 || c.sup > 0;
            if (!strong && !last) continue;
            if (try_join(uf, images, node_image, c.a, c.b)) ++merged;
            else st.conflicts += c.count;
        }
        if (last) break;
        if (merged == 0) {
            std::vector<CompPair> rest = component_pairs(zero, uf);
            stable_desc(rest, [](const CompPair& x) { return x.count; });
            for (const CompPair& c : rest)
                if (!try_join(uf, images, node_image, c.a, c.b)) st.conflicts += c.count;
            break;
        }
    }
}

}  // namespace

// The displacement flags of every pair's deduplicated matches (pairs with at
// least 8 and a usable extent; the others keep 0), on the device.
int Ctx::displacement_flags(const std::vector<std::vector<std::array<uint64_t, 2>>>& valid,
                            const std::vector<std::pair<const double*, const double*>>& kps,
                            std::vector<std::vector<uint8_t>>& flags) {
    // per pair: points, extent, grid cell and sorted grid keys (independent per pair)
    struct Prep {
        bool use = false;
        double cell = 0.0, tol = 0.0;
        std::vector<double4> q;
        std::vector<CellKey> kk;
    };
    std::vector<Prep> prep(valid.size());
    flags.assign(valid.size(), {});
    parallel_for(valid.size(), [&](size_t p) {
        const auto& m = valid[p];
        const size_t n = m.size();
        flags[p].assign(n, 0);
        if (n < 8) return;  // tracks.rs:527-529
        const double* ka = kps[p].first;
        const double* kb = kps[p].second;
        Prep& P = prep[p];
        std::vector<double4>& q = P.q;
        q.resize(n);
        double max_x = 0.0, max_y = 0.0, min_x = DBL_MAX, min_y = DBL_MAX;
        for (size_t k = 0; k < n; ++k) {
            const uint64_t fa = m[k][0], fb = m[k][1];
            q[k] = make_double4(ka[2 * fa], ka[2 * fa + 1], kb[2 * fb] - ka[2 * fa],
                                kb[2 * fb + 1] - ka[2 * fa + 1]);
            max_x = std::fmax(max_x, q[k].x);
            max_y = std::fmax(max_y, q[k].y);
            min_x = std::fmin(min_x, q[k].x);
            min_y = std::fmin(min_y, q[k].y);
        }
        const double size = std::fmax(max_x / kRefW, max_y / kRefH);
        if (size <= 0.0 || !std::isfinite(size)) return;
        const double tol = kDispTolerance * size;
        // 5 % / 95 % order statistics of each coordinate (tracks.rs:552-558)
        auto q_range = [&](bool y) {
            std::vector<double> v(n);
            for (size_t k = 0; k < n; ++k) v[k] = y ? q[k].y : q[k].x;
            const size_t lo = n / 20, hi = n - 1 - n / 20;
            std::nth_element(v.begin(), v.begin() + lo, v.end());
            const double a = v[lo];
            std::nth_element(v.begin(), v.begin() + hi, v.end());
            return std::fmax(v[hi] - a, 0.0);
        };
        const double full = std::fmax((max_x - min_x) * (max_y - min_y), 1e-12);
        const double quant = std::fmax(q_range(false) * q_range(true) / 0.81, 1e-12);
        const double area = full > 16.0 * quant ? quant : full;
        const double cell =
            std::fmax(std::fmax(std::sqrt(area * 8.0 / (double)n), size * 4.0), 1e-9);
        P.kk.resize(n);
        for (size_t k = 0; k < n; ++k)
            P.kk[k] = {sat_i64(std::floor(q[k].x / cell)), sat_i64(std::floor(q[k].y / cell)),
                       (uint32_t)k, 0u};
        std::sort(P.kk.begin(), P.kk.end(), [](const CellKey& a, const CellKey& b) {
            return std::tie(a.kx, a.ky, a.i) < std::tie(b.kx, b.ky, b.i);
        });
        P.cell = cell;
        P.tol = tol;
        P.use = true;
    });
    std::vector<DispPair> pairs;
    std::vector<std::pair<size_t, uint32_t>> where;  // pair index, first point
    size_t total = 0;
    for (size_t p = 0; p < prep.size(); ++p)
        if (prep[p].use) {
            if (total + prep[p].q.size() >= 0xffffffffull)
                return fail(SKH_EINVAL, "skh_build_tracks: too many matches");
            pairs.push_back({(uint32_t)total, (uint32_t)prep[p].q.size(), prep[p].cell,
                             prep[p].tol});
            where.emplace_back(p, (uint32_t)total);
            total += prep[p].q.size();
        }
    std::vector<uint32_t> point_pair(total);
    std::vector<double4> pts(total);
    std::vector<CellKey> keys(total);
    parallel_for(pairs.size(), [&](size_t d) {
        const Prep& P = prep[where[d].first];
        const size_t p0 = pairs[d].p0, n = P.q.size();
        std::copy(P.q.begin(), P.q.end(), pts.begin() + p0);
        std::copy(P.kk.begin(), P.kk.end(), keys.begin() + p0);
        std::fill(point_pair.begin() + p0, point_pair.begin() + p0 + n, (uint32_t)d);
    });
    if (pts.empty()) return SKH_OK;
    const size_t np = pts.size();
    int rc;
    if ((rc = ensure(t_pairs, pairs.size() * sizeof(DispPair), "track pairs")) ||
        (rc = ensure(t_point_pair, np * sizeof(uint32_t), "track point pairs")) ||
        (rc = ensure(t_pts, np * sizeof(double4), "track points")) ||
        (rc = ensure(t_keys, np * sizeof(CellKey), "track keys")) ||
        (rc = ensure(t_flags, np, "track flags")))
        return rc;
    if ((rc = hip(hipMemcpyAsync(t_pairs.ptr, pairs.data(), pairs.size() * sizeof(DispPair),
                                 hipMemcpyHostToDevice, stream), "copy track pairs")) ||
        (rc = hip(hipMemcpyAsync(t_point_pair.ptr, point_pair.data(), np * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, stream), "copy track point pairs")) ||
        (rc = hip(hipMemcpyAsync(t_pts.ptr, pts.data(), np * sizeof(double4),
                                 hipMemcpyHostToDevice, stream), "copy track points")) ||
        (rc = hip(hipMemcpyAsync(t_keys.ptr, keys.data(), np * sizeof(CellKey),
                                 hipMemcpyHostToDevice, stream), "copy track keys")))
        return rc;
    hipLaunchKernelGGL(displacement_kernel, dim3((uint32_t)((np + 63) / 64)), dim3(64), 0, stream,
                       t_pairs.as<DispPair>(), t_point_pair.as<uint32_t>(), t_pts.as<double4>(),
                       t_keys.as<CellKey>(), (uint32_t)np, kDispFloor, kDispCap,
                       t_flags.as<uint8_t>());
    if ((rc = hip(hipGetLastError(), "launch displacement_kernel"))) return rc;
    std::vector<uint8_t> h(np);
    if ((rc = hip(hipMemcpyAsync(h.data(), t_flags.ptr, np, hipMemcpyDeviceToHost, stream),
                  "copy track flags")) ||
        (rc = hip(hipStreamSynchronize(stream), "sync")))
        return rc;
    for (const auto& w : where)
        std::copy(h.begin() + w.second, h.begin() + w.second + flags[w.first].size(),
                  flags[w.first].begin());
    return SKH_OK;
}

}  // namespace skh

using namespace skh;

extern "C" void skh_track_cfg_default(skh_track_cfg* cfg) {
    if (!cfg) return;
    cfg->policy = SKH_TRACKS_SPLIT;  // TrackConfig::default (tracks.rs:118-125)
    cfg->pad = 0;
    cfg->min_length = 3;
    cfg->max_tracks = 100000;  // MAX_TRACKS
}

extern "C" int skh_build_tracks(skh_ctx* ctx, size_t n_pairs, const uint64_t* image_a,
                                const uint64_t* image_b, const uint64_t* m_off,
                                const uint64_t* m_ab, size_t n_images, const uint64_t* kp_off,
                                const double* kp, const skh_track_cfg* cfg, uint64_t* track_off,
                                size_t track_cap, uint64_t* obs, size_t obs_cap,
                                size_t* n_tracks, skh_track_stats* stats) {
    if (!ctx) return SKH_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (!cfg || !kp_off || !n_tracks || !stats || !track_off || track_cap < 1 ||
        (n_pairs && (!image_a || !image_b || !m_off)) ||
        (cfg->policy != SKH_TRACKS_DROP && cfg->policy != SKH_TRACKS_SPLIT))
        return ctx->fail(SKH_EINVAL, "skh_build_tracks: invalid argument");
    if (n_images > UINT32_MAX) return ctx->fail(SKH_EINVAL, "skh_build_tracks: n_images");
    const size_t n = (size_t)kp_off[n_images];
    if ((n && !kp) || (n_pairs && m_off[n_pairs] && !m_ab))
        return ctx->fail(SKH_EINVAL, "skh_build_tracks: invalid argument");
    for (size_t i = 0; i < n_images; ++i)
        if (kp_off[i + 1] < kp_off[i]) return ctx->fail(SKH_EINVAL, "skh_build_tracks: kp_off");
    for (size_t p = 0; p < n_pairs; ++p)
        if (m_off[p + 1] < m_off[p]) return ctx->fail(SKH_EINVAL, "skh_build_tracks: m_off");
    if (hipSetDevice(ctx->device) != hipSuccess) return SKH_ENODEV;
    skh_track_stats st{};
    std::vector<size_t> node_image(n);
    for (size_t i = 0; i < n_images; ++i)
        for (uint64_t x = kp_off[i]; x < kp_off[i + 1]; ++x) node_image[x] = i;
    auto n_kp = [&](uint64_t img) { return kp_off[img + 1] - kp_off[img]; };
    // tracks.rs:207-230: valid matches in the lower image's direction, sorted, deduplicated
    std::vector<std::vector<std::array<uint64_t, 2>>> valid(n_pairs);
    std::vector<std::pair<const double*, const double*>> kps(n_pairs, {nullptr, nullptr});
    std::vector<size_t> invalid(n_pairs, 0);
    parallel_for(n_pairs, [&](size_t p) {
        const uint64_t ia = image_a[p], ib = image_b[p];
        auto& v = valid[p];
        v.reserve(m_off[p + 1] - m_off[p]);
        for (uint64_t k = m_off[p]; k < m_off[p + 1]; ++k) {
            const uint64_t fa = m_ab[2 * k], fb = m_ab[2 * k + 1];
            if (ia == ib || ia >= n_images || ib >= n_images || fa >= n_kp(ia) || fb >= n_kp(ib)) {
                invalid[p]++;
                continue;
            }
            v.push_back(ia > ib ? std::array<uint64_t, 2>{fb, fa} : std::array<uint64_t, 2>{fa, fb});
        }
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        if (!v.empty()) {
            const uint64_t lo = std::min(ia, ib), hi = std::max(ia, ib);
            kps[p] = {kp + 2 * kp_off[lo], kp + 2 * kp_off[hi]};
        }
    });
    for (size_t p = 0; p < n_pairs; ++p) st.invalid_matches += invalid[p];
    std::vector<std::vector<uint8_t>> flags;
    if (cfg->policy == SKH_TRACKS_SPLIT) {
        const int rc = ctx->displacement_flags(valid, kps, flags);
        if (rc) return rc;
    }
    std::vector<Edge> edges;
    for (size_t p = 0; p < n_pairs; ++p) {
        if (valid[p].empty()) continue;
        const uint64_t lo = std::min(image_a[p], image_b[p]), hi = std::max(image_a[p], image_b[p]);
        for (size_t k = 0; k < valid[p].size(); ++k) {
            if (cfg->policy == SKH_TRACKS_SPLIT && flags[p][k]) {
                st.inconsistent++;
                continue;
            }
            const size_t u = kp_off[lo] + valid[p][k][0], v = kp_off[hi] + valid[p][k][1];
            edges.emplace_back(std::min(u, v), std::max(u, v));
        }
    }
    sort_pairs(edges, n);
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    st.edges = edges.size();
    UnionFind uf(n);
    if (cfg->policy == SKH_TRACKS_DROP) {
        for (const Edge& e : edges) {
            const size_t ru = uf.find(e.first), rv = uf.find(e.second);
            if (ru != rv) uf.link(ru, rv);
        }
    } else {
        split_policy(edges, n, node_image, uf, st);
    }
    // components of the remaining edges' end points (tracks.rs:467-508): the
    // distinct (root, node) pairs in ascending order, by a counting sort on
    // the root over the nodes taken in ascending order
    std::vector<std::pair<size_t, size_t>> members;  // (root, node)
    {
        std::vector<size_t> root(n, SIZE_MAX), at(n + 1, 0);
        for (const Edge& e : edges) root[e.first] = root[e.second] = 0;
        size_t used = 0;
        for (size_t x = 0; x < n; ++x)
            if (root[x] != SIZE_MAX) {
                root[x] = uf.find(x);
                at[root[x] + 1]++;
                ++used;
            }
        for (size_t r = 0; r < n; ++r) at[r + 1] += at[r];
        members.resize(used);
        for (size_t x = 0; x < n; ++x)
            if (root[x] != SIZE_MAX) members[at[root[x]]++] = {root[x], x};
    }
    const size_t min_len = std::max<uint64_t>(cfg->min_length, 2);
    struct Tr {
        size_t first, len;
    };
    std::vector<Tr> tracks;
    std::vector<size_t> nodes;  // accepted tracks' nodes, track after track
    for (size_t g = 0; g < members.size();) {
        size_t h = g;
        while (h < members.size() && members[h].first == members[g].first) ++h;
        bool conflict = false;
        for (size_t k = g + 1; k < h; ++k)
            conflict |= node_image[members[k].second] == node_image[members[k - 1].second];
        if (conflict) {
            st.conflicts++;
        } else if (h - g < min_len) {
            st.too_short++;
        } else {
            tracks.push_back({nodes.size(), h - g});
            for (size_t k = g; k < h; ++k) nodes.push_back(members[k].second);
        }
        g = h;
    }
    // select_tracks (tracks.rs:786-799)
    auto first_key = [&](const Tr& t) {
        const size_t x = nodes[t.first];
        return std::make_pair(node_image[x], x - (size_t)kp_off[node_image[x]]);
    };
    auto span = [&](const Tr& t) {
        return node_image[nodes[t.first + t.len - 1]] - node_image[nodes[t.first]];
    };
    const size_t before = tracks.size();
    if (tracks.size() > cfg->max_tracks) {
        std::sort(tracks.begin(), tracks.end(), [&](const Tr& a, const Tr& b) {
            if (a.len != b.len) return a.len > b.len;
            if (span(a) != span(b)) return span(a) > span(b);
            return first_key(a) < first_key(b);
        });
        tracks.resize(cfg->max_tracks);
    }
    std::sort(tracks.begin(), tracks.end(),
              [&](const Tr& a, const Tr& b) { return first_key(a) < first_key(b); });
    st.truncated = before - tracks.size();
    st.tracks = tracks.size();
    size_t total = 0;
    for (const Tr& t : tracks) total += t.len;
    *stats = st;
    *n_tracks = tracks.size();
    if (tracks.size() + 1 > track_cap || total > obs_cap || (total && !obs))
        return ctx->fail(SKH_EINVAL, "skh_build_tracks: %zu tracks, %zu observations exceed the "
                                     "capacity", tracks.size(), total);
    size_t w = 0;
    track_off[0] = 0;
    for (size_t t = 0; t < tracks.size(); ++t) {
        for (size_t k = 0; k < tracks[t].len; ++k) {
            const size_t x = nodes[tracks[t].first + k];
            obs[2 * w] = node_image[x];
            obs[2 * w + 1] = x - kp_off[node_image[x]];
            ++w;
        }
        track_off[t + 1] = w;
    }
    return SKH_OK;
}
