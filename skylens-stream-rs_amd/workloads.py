"""Seeded synthetic workloads for the matching path (numpy; scene pair lists
come from the library's candidate_pairs).

Descriptors are unit 128-d f32 rows; keypoints are f32 pixel positions, one
per feature, in the bench's raw pixel convention (benches/pipeline.rs:276-283
gathers `kp.x as f64` without +0.5).

* ``single_pair()`` -- the single CPU pair: A 800 unit Gaussian rows; B 480
  noisy copies (sigma 0.05, renormalised) of distinct random A rows plus 320
  fresh rows, permuted; ratio 0.8, mutual. Keypoints are tied to the rows:
  copies see their source's 3D point (two pinholes fx=fy=300, 320x180, 5 m
  baseline, 15 deg yaw, sigma 0.5 px), 40 % of them are uniform outliers, so
  the matches handed to ransac_fundamental carry the geometry.
* ``scene(...)`` -- a multi-view scene: 3 cameras x P positions (view index
  3p + c, matching synth.rs:284-305), per-view feature counts drawn by the
  caller's rule, 60 % noisy copies (sigma 0.05) of a shared pool of 20k unit
  vectors and 40 % fresh rows. Every pool vector owns a 3D point; a view only
  copies pool vectors whose point it sees, and a copy's keypoint is that
  point's projection (960x540, fx=fy=750, sigma 0.5 px) unless the view's
  outlier draw replaces it with a uniform position. Each view has its own
  outlier rate o_v, so a true match survives with (1 - o_a)(1 - o_b).
  - ``full_scene()``: 240 views, M ~ U[1500, 4096], pair outlier ratio in
    [20 %, 60 %] (o_v ~ U[0.106, 0.368]); 3663 candidate pairs.
  - ``uneven_scene()``: 264 views, M log-uniform in [200, 8000], pair inlier
    ratio in [10 %, 90 %] (o_v ~ U[0.051, 0.684]); the first 4000 of the 4047
    candidate pairs (48P - 177 has no P giving exactly 4000).
  - ``fixed_scene()``: 240 views with 8192 features each.
* ``textured_image(w, h, seed)`` -- a grey image for the detector: random
  Gaussian blobs (radius 1.5-6.5 px, one per 400 px^2) over a gentle gradient.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

DESC = 128
POOL = 20000


def unit_rows(rng: np.random.Generator, n: int) -> np.ndarray:
    x = rng.standard_normal((n, DESC), dtype=np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    return x.astype(np.float32)


def noisy_copies(rng: np.random.Generator, src: np.ndarray, sigma: float) -> np.ndarray:
    y = src + np.float32(sigma) * rng.standard_normal(src.shape, dtype=np.float32)
    y /= np.linalg.norm(y, axis=1, keepdims=True)
    return y.astype(np.float32)


def scene_pair_descriptors(seed: int, na: int, nb: int, pool: int = 4000) -> Tuple[np.ndarray, np.ndarray]:
    """One pair drawn like the full scene (60 % pool copies, 40 % fresh)."""
    rng = np.random.default_rng(seed)
    P = unit_rows(rng, pool)

    def view(n):
        k = int(round(0.6 * n))
        out = np.empty((n, DESC), np.float32)
        out[:k] = noisy_copies(rng, P[rng.integers(0, pool, k)], 0.05)
        out[k:] = unit_rows(rng, n - k)
        return out[rng.permutation(n)]

    return view(na), view(nb)


# --------------------------------------------------------------------------
# cameras
# --------------------------------------------------------------------------

def rot_y(deg: float) -> np.ndarray:
    t = np.deg2rad(deg)
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


@dataclass
class Camera:
    R: np.ndarray  # world -> camera rotation
    C: np.ndarray  # centre
    f: float
    cx: float
    cy: float
    w: int
    h: int

    def project(self, X: np.ndarray):
        """(n, 3) world points -> (n, 2) pixels and a visibility mask."""
        xc = (X - self.C) @ self.R.T
        z = xc[:, 2]
        with np.errstate(divide="ignore", invalid="ignore"):
            u = self.f * xc[:, 0] / z + self.cx
            v = self.f * xc[:, 1] / z + self.cy
        vis = (z > 0) & (u >= 0) & (v >= 0) & (u < self.w) & (v < self.h)
        return np.stack([u, v], 1), vis


# --------------------------------------------------------------------------
# single CPU pair
# --------------------------------------------------------------------------

@dataclass
class SinglePair:
    a: np.ndarray         # (800, 128) f32
    b: np.ndarray         # (800, 128) f32
    b_source: np.ndarray  # A row each B row copies, -1 for fresh
    kp_a: np.ndarray      # (800, 2) f32 keypoints of A's features (image 1)
    kp_b: np.ndarray      # (800, 2) f32 keypoints of B's features (image 2)
    b_inlier: np.ndarray  # B row's keypoint is its source's 3D point seen in image 2

    def handoff(self, matches, px_offset: float = 0.0):
        """x1, x2 (n, 2) f64 of a match list (benches/pipeline.rs:276-283)."""
        m = np.asarray(matches, np.int64).reshape(-1, 2)
        x1 = self.kp_a[m[:, 0]].astype(np.float64) + px_offset
        x2 = self.kp_b[m[:, 1]].astype(np.float64) + px_offset
        return x1, x2


def single_pair(seed: int = 0) -> SinglePair:
    """A 800 unit Gaussian rows; B 480 noisy copies (sigma 0.05) of distinct A rows
    plus 320 fresh rows, permuted. Keypoints are tied to the descriptors: 600 3D
    points (uniform in a 100 x 100 x 20 m box, visible in both 320x180 pinholes,
    fx = fy = 300, 5 m baseline, 15 deg yaw) belong to the 480 copied A rows and
    120 more; a copy's keypoint in image 2 is its point's projection (sigma 0.5
    px) unless it is one of the 40 % outliers, drawn uniformly over the image,
    like the fresh rows and the A rows without a point. So the correspondences
    that ratio_match hands to ransac_fundamental are ~60 % inliers."""
    rng = np.random.default_rng(seed)
    na = 800
    a = unit_rows(rng, na)
    src = rng.choice(na, 480, replace=False)
    b = np.concatenate([noisy_copies(rng, a[src], 0.05), unit_rows(rng, 320)])
    b_source = np.concatenate([src, -np.ones(320, np.int64)])

    w, h, f = 320, 180, 300.0
    c1 = Camera(np.eye(3), np.zeros(3), f, 160.0, 90.0, w, h)
    c2 = Camera(rot_y(15.0), np.array([5.0, 0.0, 0.0]), f, 160.0, 90.0, w, h)
    pts = []
    while sum(len(p) for p in pts) < 600:
        X = np.stack([rng.uniform(-50, 50, 4000), rng.uniform(-50, 50, 4000),
                      rng.uniform(60, 80, 4000)], 1)
        p1, v1 = c1.project(X)
        p2, v2 = c2.project(X)
        ok = v1 & v2
        pts.append(np.concatenate([p1[ok], p2[ok]], 1))
    P = np.concatenate(pts)[:600]
    others = np.setdiff1d(np.arange(na), src)
    owners = np.concatenate([src, rng.choice(others, 120, replace=False)])
    point_of = -np.ones(na, np.int64)
    point_of[owners] = np.arange(600)

    def uniform(n):
        return np.stack([rng.uniform(0, w, n), rng.uniform(0, h, n)], 1)

    kp_a = uniform(na)
    kp_a[owners] = P[:, :2] + 0.5 * rng.standard_normal((600, 2))
    kp_b = uniform(800)
    kp_b[:480] = P[point_of[src], 2:] + 0.5 * rng.standard_normal((480, 2))
    inl = np.zeros(800, bool)
    inl[:480] = rng.random(480) >= 0.4
    kp_b[:480][~inl[:480]] = uniform(int((~inl[:480]).sum()))
    perm = rng.permutation(800)

    def img(k):
        return np.clip(k, 0.0, [w - 1e-3, h - 1e-3]).astype(np.float32)

    return SinglePair(a, b[perm], b_source[perm], img(kp_a), img(kp_b[perm]), inl[perm])


def two_view_correspondences(n: int, sigma: float, outlier_frac: float, seed: int):
    """n pixel correspondences (f64) between two 960x540 pinholes (fx = fy = 750,
    5 m baseline, 15 deg yaw): (1 - outlier_frac) n projections of 3D points
    visible in both (sigma px noise), the rest uniform over the images, shuffled
    -- the shape of the reference's 4000-correspondence timing case
    (crates/core/src/matching.rs:2133-2149)."""
    rng = np.random.default_rng(seed)
    W, H, F = 960, 540, 750.0
    c1 = Camera(np.eye(3), np.zeros(3), F, W / 2, H / 2, W, H)
    c2 = Camera(rot_y(15.0), np.array([5.0, 0.0, 0.0]), F, W / 2, H / 2, W, H)
    k = int(round((1.0 - outlier_frac) * n))
    pts = []
    while sum(len(p) for p in pts) < k:
        X = np.stack([rng.uniform(-40, 40, 4 * n), rng.uniform(-25, 25, 4 * n),
                      rng.uniform(50, 90, 4 * n)], 1)
        p1, v1 = c1.project(X)
        p2, v2 = c2.project(X)
        ok = v1 & v2
        pts.append(np.concatenate([p1[ok], p2[ok]], 1))
    P = np.concatenate(pts)[:k]
    x1 = np.concatenate([P[:, :2] + sigma * rng.standard_normal((k, 2)),
                         np.stack([rng.uniform(0, W, n - k), rng.uniform(0, H, n - k)], 1)])
    x2 = np.concatenate([P[:, 2:] + sigma * rng.standard_normal((k, 2)),
                         np.stack([rng.uniform(0, W, n - k), rng.uniform(0, H, n - k)], 1)])
    perm = rng.permutation(n)
    return np.ascontiguousarray(x1[perm]), np.ascontiguousarray(x2[perm])


# --------------------------------------------------------------------------
# multi-view scenes
# --------------------------------------------------------------------------

@dataclass
class Scene:
    views: List[Tuple[int, int]]      # (camera, position), index 3p + c
    counts: np.ndarray                # features per view
    offsets: np.ndarray               # CSR offsets into desc/kp (len V + 1)
    desc: np.ndarray                  # (total, 128) f32
    kp: np.ndarray                    # (total, 2) f32 pixel positions
    pool_id: np.ndarray               # (total,) pool vector copied, -1 fresh
    outlier: np.ndarray               # (total,) keypoint replaced by a uniform draw
    pairs: np.ndarray                 # (n_pairs, 2) view indices
    width: int = 960
    height: int = 540

    def view_desc(self, v: int) -> np.ndarray:
        return self.desc[self.offsets[v]:self.offsets[v + 1]]

    def view_kp(self, v: int) -> np.ndarray:
        return self.kp[self.offsets[v]:self.offsets[v + 1]]


def candidate_pair_array(views, temporal=5, cross=4, pow2_max=16) -> np.ndarray:
    """The scene's pair list from the library's candidate_pairs (matching.rs:18-40,
    host code behind the C-ABI), as an (n, 2) int64 array."""
    import skylens_hip

    pairs = skylens_hip.candidate_pairs(views, temporal, cross, pow2_max)
    return np.asarray(pairs, np.int64).reshape(-1, 2)


def _candidate_pairs(views, temporal=5, cross=4, pow2_max=16) -> np.ndarray:
    """matching.rs:18-40, vectorised (i ascending, then j): an independent
    restatement the tests compare the library's list with."""
    v = np.asarray(views, np.int64)
    n = len(v)
    out = []
    for i in range(n):
        j = np.arange(i + 1, n)
        ca, pa = v[i]
        cb, pb = v[j, 0], v[j, 1]
        d = np.abs(pb - pa)
        pow2 = (d > 0) & ((d & (d - 1)) == 0)
        same = (d >= 1) & ((d <= temporal) | (pow2 & (d <= pow2_max)))
        ok = np.where(cb == ca, same, d <= cross)
        out.append(np.stack([np.full(ok.sum(), i), j[ok]], 1))
    return np.concatenate(out).astype(np.int64) if out else np.zeros((0, 2), np.int64)


def scene(seed: int, positions: int, counts: np.ndarray, outlier_lo: float, outlier_hi: float,
          max_pairs: Optional[int] = None, pool_size: int = POOL) -> Scene:
    rng = np.random.default_rng(seed)
    W, H, F = 960, 540, 750.0
    pool = unit_rows(rng, pool_size)
    X = np.stack([rng.uniform(-40, positions + 40, pool_size), rng.uniform(-30, 30, pool_size),
                  rng.uniform(60, 80, pool_size)], 1)
    views = [(c, p) for p in range(positions) for c in range(3)]
    yaw = {0: 0.0, 1: 15.0, 2: -15.0}
    V = len(views)
    assert len(counts) == V
    offsets = np.zeros(V + 1, np.int64)
    offsets[1:] = np.cumsum(counts)
    total = int(offsets[-1])
    desc = np.empty((total, DESC), np.float32)
    kp = np.empty((total, 2), np.float32)
    pool_id = np.full(total, -1, np.int64)
    outl = np.zeros(total, bool)
    for v, (c, p) in enumerate(views):
        m = int(counts[v])
        cam = Camera(rot_y(yaw[c]), np.array([float(p), 0.0, 0.0]), F, W / 2, H / 2, W, H)
        px, vis = cam.project(X)
        visible = np.flatnonzero(vis)
        o_v = rng.uniform(outlier_lo, outlier_hi)
        k = int(round(0.6 * m))
        q = visible[rng.integers(0, len(visible), k)]
        d = np.empty((m, DESC), np.float32)
        d[:k] = noisy_copies(rng, pool[q], 0.05)
        d[k:] = unit_rows(rng, m - k)
        key = np.empty((m, 2))
        key[:k] = px[q] + 0.5 * rng.standard_normal((k, 2))
        key[k:] = np.stack([rng.uniform(0, W, m - k), rng.uniform(0, H, m - k)], 1)
        rep = np.zeros(m, bool)
        rep[:k] = rng.random(k) < o_v
        key[rep] = np.stack([rng.uniform(0, W, rep.sum()), rng.uniform(0, H, rep.sum())], 1)
        pid = np.full(m, -1, np.int64)
        pid[:k] = q
        perm = rng.permutation(m)
        s = slice(offsets[v], offsets[v + 1])
        desc[s] = d[perm]
        kp[s] = np.clip(key[perm], 0.0, [W - 1e-3, H - 1e-3]).astype(np.float32)
        pool_id[s] = pid[perm]
        outl[s] = rep[perm]
    pairs = candidate_pair_array(views)
    if max_pairs is not None:
        pairs = pairs[:max_pairs]
    return Scene(views, np.asarray(counts, np.int64), offsets, desc, kp, pool_id, outl, pairs,
                 W, H)


def full_scene(seed: int = 1, positions: int = 80) -> Scene:
    rng = np.random.default_rng(seed + 1000)
    counts = rng.integers(1500, 4097, 3 * positions)
    return scene(seed, positions, counts, 0.106, 0.368)


def uneven_scene(seed: int = 4, positions: int = 88, n_pairs: int = 4000) -> Scene:
    rng = np.random.default_rng(seed + 1000)
    counts = np.exp(rng.uniform(np.log(200), np.log(8000), 3 * positions)).astype(np.int64)
    return scene(seed, positions, counts, 0.051, 0.684, max_pairs=n_pairs)


def fixed_scene(seed: int = 3, positions: int = 80, features: int = 8192) -> Scene:
    return scene(seed, positions, np.full(3 * positions, features), 0.106, 0.368)


def textured_image(w: int, h: int, seed: int, n_blobs: Optional[int] = None) -> np.ndarray:
    """A (h, w) float32 detector test image: random Gaussian blobs of radius
    1.5-6.5 px and contrast +-0.25 at one per 400 px^2 over a 0.4 -> 0.6
    horizontal gradient (the reference renders a synthetic drone scene,
    synth.rs, which is outside this path)."""
    rng = np.random.default_rng(seed)
    n = n_blobs if n_blobs is not None else max(1, (w * h) // 400)
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    v = (np.float32(0.4) + np.float32(0.2) * x / np.float32(max(w, 1))).astype(np.float32)
    bx = rng.uniform(0, w, n).astype(np.float32)
    by = rng.uniform(0, h, n).astype(np.float32)
    bs = rng.uniform(1.5, 6.5, n).astype(np.float32)
    ba = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    for i in range(n):
        r = int(4 * bs[i]) + 1
        x0, x1 = max(0, int(bx[i]) - r), min(w, int(bx[i]) + r + 1)
        y0, y1 = max(0, int(by[i]) - r), min(h, int(by[i]) + r + 1)
        if x0 >= x1 or y0 >= y1:
            continue
        d2 = (x[y0:y1, x0:x1] - bx[i]) ** 2 + (y[y0:y1, x0:x1] - by[i]) ** 2
        v[y0:y1, x0:x1] += np.float32(0.5) * ba[i] * np.exp(-d2 / (np.float32(2.0) * bs[i] * bs[i]))
    return v


@dataclass
class TrackScene:
    """Verified pairwise matches of a sequence of views of one planar scene:
    keypoints[i] (n_i, 2) pixels of image i, gt[i][f] the scene point of
    feature f (negative for clutter), pairs (image_a, image_b, (m, 2) matches)."""
    keypoints: List[np.ndarray]
    gt: List[np.ndarray]
    pairs: List[Tuple[int, int, np.ndarray]]
    inliers: int
    outliers: int


def track_scene(seed: int = 0, images: int = 12, points: int = 2500, keep: float = 1.0,
                outlier_per_mille: int = 10, temporal: int = 5, width: float = 960.0,
                height: float = 540.0) -> TrackScene:
    """Views sliding along a strip of scene points (80 px per image, a small
    rotation and scale per view, 0.3 px noise), 3 % clutter features per
    image, pairs (i, j) with 1 <= j - i <= temporal (every other pair stored
    in the (j, i) direction), each shared point matched with probability
    ``keep`` plus ``outlier_per_mille`` random wrong matches per 1000 true ones."""
    rng = np.random.default_rng(seed)
    span = width + 80.0 * images
    world = np.stack([rng.uniform(0.0, span, points), rng.uniform(0.0, height, points)], 1)
    kps, gts = [], []
    for i in range(images):
        th, sc = 0.004 * (i % 5) - 0.008, 1.0 + 0.006 * (i % 3)
        c, s = np.cos(th), np.sin(th)
        local = (world - np.array([80.0 * i, 0.0])) * sc
        px = np.stack([c * local[:, 0] - s * local[:, 1], s * local[:, 0] + c * local[:, 1]], 1)
        px = px + rng.normal(0.0, 0.3, px.shape)
        vis = np.nonzero((px[:, 0] >= 0) & (px[:, 0] < width) & (px[:, 1] >= 0)
                         & (px[:, 1] < height))[0]
        n_cl = int(0.03 * len(vis))
        clutter = np.stack([rng.uniform(0, width, n_cl), rng.uniform(0, height, n_cl)], 1)
        feats = np.concatenate([px[vis], clutter])
        g = np.concatenate([vis, -1 - np.arange(n_cl) - 1000 * i])
        perm = rng.permutation(len(feats))
        kps.append(np.ascontiguousarray(feats[perm]))
        gts.append(g[perm])
    pairs, n_in, n_out = [], 0, 0
    for i in range(images):
        for j in range(i + 1, min(images, i + temporal + 1)):
            fa_of = {int(p): f for f, p in enumerate(gts[i]) if p >= 0}
            fb_of = {int(p): f for f, p in enumerate(gts[j]) if p >= 0}
            shared = sorted(set(fa_of) & set(fb_of))
            m = [(fa_of[p], fb_of[p]) for p in shared if rng.random() < keep]
            n_in += len(m)
            n_bad = len(m) * outlier_per_mille // 1000
            for _ in range(n_bad):
                fa, fb = int(rng.integers(len(gts[i]))), int(rng.integers(len(gts[j])))
                if gts[i][fa] != gts[j][fb]:
                    m.append((fa, fb))
                    n_out += 1
            m = np.array(m, np.int64).reshape(-1, 2)
            m = m[rng.permutation(len(m))]
            if (i + j) % 2:
                pairs.append((j, i, np.ascontiguousarray(m[:, ::-1])))
            else:
                pairs.append((i, j, m))
    return TrackScene(kps, gts, pairs, n_in, n_out)
