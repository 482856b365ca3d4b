"""The detector oracle (oracle/features_oracle.c) against the reference's own
detector tests (crates/core/src/features.rs:819-1335), restated, plus the
checks that pin its parts the reference's tests leave open:

- the f32 transcendentals it evaluates per pixel equal the correctly rounded
  values (Python's double-precision math rounded to f32) and glibc's f32
  functions within one unit in the last place;
- gaussian_blur equals a direct numpy restatement (edge-clamped separable
  convolution in tap order) bit for bit, and its taps equal the closed form;
- upsample2 / downsample equal their definitions;
- the descriptor of a keypoint equals orc_describe on the same blurred layer
  (the public describe() of features.rs:509-519 and the detector's own path
  agree), and dominant_orientations follow a rotated ramp.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob


def f32(v):
    return np.float32(v)


def ulp_diff(a, b):
    ia = np.array([a], np.float32).view(np.int32)[0]
    ib = np.array([b], np.float32).view(np.int32)[0]
    return abs(int(ia) - int(ib))


# ---- the per-pixel f32 transcendentals -------------------------------------

@pytest.fixture(scope="module")
def libm():
    import ctypes
    m = ctypes.CDLL("libm.so.6")
    for n, k in (("expf", 1), ("sinf", 1), ("cosf", 1), ("atan2f", 2), ("powf", 2)):
        getattr(m, n).restype = ctypes.c_float
        getattr(m, n).argtypes = [ctypes.c_float] * k
    return m


def test_expf_correctly_rounded_and_within_one_ulp_of_glibc(libm):
    rng = np.random.default_rng(11)
    xs = np.concatenate([rng.uniform(-40, 2, 20000), rng.uniform(-1e-3, 1e-3, 2000),
                         [0.0, -0.0, -87.0, -103.0, 1.0, -1.0]]).astype(np.float32)
    glibc_off = 0
    for x in xs:
        got = ob.mathf("expf", float(x))
        assert got == f32(math.exp(float(x))), x
        d = ulp_diff(got, libm.expf(float(x)))
        assert d <= 1
        glibc_off += d
    assert glibc_off < 0.01 * len(xs)


def test_atan2f_correctly_rounded_and_within_one_ulp_of_glibc(libm):
    rng = np.random.default_rng(12)
    ys = rng.normal(size=20000).astype(np.float32)
    xs = rng.normal(size=20000).astype(np.float32)
    for y, x in zip(ys, xs):
        got = ob.mathf("atan2f", float(y), float(x))
        assert got == f32(math.atan2(float(y), float(x))), (y, x)
        assert ulp_diff(got, libm.atan2f(float(y), float(x))) <= 1
    # IEEE zeros and axes
    for y, x in [(0.0, 0.0), (-0.0, 0.0), (0.0, -0.0), (-0.0, -0.0), (1.0, 0.0), (-1.0, 0.0),
                 (0.0, -2.0), (3.0, 3.0), (-3.0, -3.0)]:
        got = ob.mathf("atan2f", y, x)
        want = f32(math.atan2(y, x))
        assert got == want and math.copysign(1.0, got) == math.copysign(1.0, want), (y, x)


def test_sincosf_correctly_rounded_and_within_one_ulp_of_glibc(libm):
    rng = np.random.default_rng(13)
    a = np.concatenate([rng.uniform(0, 2 * np.pi, 20000), [0.0, np.pi / 2, np.pi, 6.2831855]])
    for v in a.astype(np.float32):
        s, c = ob.mathf("sincos", float(v))
        assert s == f32(math.sin(float(v))) and c == f32(math.cos(float(v))), v
        assert ulp_diff(s, libm.sinf(float(v))) <= 1 and ulp_diff(c, libm.cosf(float(v))) <= 1


def test_powf_correctly_rounded_on_the_detector_scale_step(libm):
    rng = np.random.default_rng(14)
    for s_ in (1, 2, 3, 4, 5):
        k = f32(libm.powf(2.0, float(f32(1.0) / f32(s_))))
        for v in rng.uniform(0.5, s_ + 0.5, 4000).astype(np.float32):
            got = ob.mathf("powf", float(k), float(v))
            assert got == f32(math.pow(float(k), float(v))), (k, v)
            assert ulp_diff(got, libm.powf(float(k), float(v))) <= 1


# ---- images ------------------------------------------------------------------

def blur_numpy(img, sigma, expf):
    """gaussian_blur (features.rs:214-256) in numpy float32, taps summed in
    order; the taps use the platform's expf, as Rust's f32::exp does."""
    img = np.asarray(img, np.float32)
    r = max(int(math.ceil(f32(3.0) * f32(sigma))), 1)
    i = np.arange(-r, r + 1)
    den = f32(2.0) * f32(sigma) * f32(sigma)
    k = np.array([expf(float(f32(-(j * j)) / den)) for j in i], np.float32)
    s = f32(0.0)
    for v in k:
        s = f32(s + v)
    k = (k / s).astype(np.float32)
    h, w = img.shape
    xs = np.clip(np.arange(w)[None, :] + i[:, None], 0, w - 1)  # (taps, w)
    tmp = np.zeros_like(img)
    for t in range(2 * r + 1):
        tmp = (tmp + k[t] * img[:, xs[t]]).astype(np.float32)
    ys = np.clip(np.arange(h)[None, :] + i[:, None], 0, h - 1)
    out = np.zeros_like(img)
    for t in range(2 * r + 1):
        out = (out + k[t] * tmp[ys[t], :]).astype(np.float32)
    return out, k


def test_blur_preserves_mean_and_constant():
    """features.rs:835-844."""
    b = ob.gaussian_blur(np.full((10, 20), 0.7, np.float32), 2.0)
    assert np.all(np.abs(b - 0.7) < 1e-5)


@pytest.mark.parametrize("sigma", [0.5, 1.2489996, 1.6, 2.0, 3.0886, 5.5])
def test_blur_equals_separable_restatement(sigma, libm):
    img = ob.texture_warped(53, 37, 3, 0.3, 1.1)
    got = ob.gaussian_blur(img, sigma)
    want, k = blur_numpy(img, sigma, libm.expf)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(ob.gaussian_taps(sigma), k)


def test_upsample_and_downsample_definitions():
    img = ob.texture_warped(31, 17, 5, 0.0, 1.0)
    up = ob.upsample2(img)
    assert up.shape == (34, 62)
    # the original pixel (x, y) sits at (2x + 0.5, 2y + 0.5): interior upsampled
    # pixels 2x and 2x+1 weigh x with 3/4 and its neighbour with 1/4
    h, w = img.shape
    for y in range(1, 2 * h - 1):
        for x in range(1, 2 * w - 1):
            sx, sy = (x + 0.5) / 2.0 - 0.5, (y + 0.5) / 2.0 - 0.5
            x0, y0 = math.floor(sx), math.floor(sy)
            fx, fy = f32(sx - x0), f32(sy - y0)
            xa, xb = min(max(x0, 0), w - 1), min(max(x0 + 1, 0), w - 1)
            ya, yb = min(max(y0, 0), h - 1), min(max(y0 + 1, 0), h - 1)
            top = f32(img[ya, xa] * (f32(1) - fx)) + f32(img[ya, xb] * fx)
            bot = f32(img[yb, xa] * (f32(1) - fx)) + f32(img[yb, xb] * fx)
            assert up[y, x] == f32(f32(top * (f32(1) - fy)) + f32(bot * fy))


# ---- the reference's detector tests, on the oracle ---------------------------

@pytest.mark.parametrize("sb,cx,cy", [(4.0, 61.0, 47.0), (8.0, 70.0, 66.0)])
def test_detects_blob_location_and_scale(sb, cx, cy):
    """features.rs:846-866."""
    kps, _ = ob.detect_and_describe(ob.blob(140, 120, cx, cy, sb), describe=False)
    assert len(kps) > 0
    x, y, sigma = kps[0][:3]
    pos_err = math.hypot(x - cx, y - cy)
    assert pos_err <= 2.0 * max(sigma / 1.6, 1.0)
    corrected = (sigma / sb) * 2 ** (1.0 / 6.0)
    assert abs(corrected - 1.0) < 0.03
    assert pos_err < 0.1


@pytest.mark.parametrize("cx,cy", [(61.3, 47.6), (60.75, 48.2)])
def test_refines_subpixel_center(cx, cy):
    """features.rs:868-878."""
    kps, _ = ob.detect_and_describe(ob.blob(140, 120, cx, cy, 4.0), describe=False)
    assert math.hypot(kps[0][0] - cx, kps[0][1] - cy) < 0.1


def angle_diff(a, b):
    d = (a - b) % (2 * math.pi)
    return min(d, 2 * math.pi - d)


def test_orientation_follows_rotation():
    """features.rs:885-912: a ramp in direction th plus a blob; the dominant
    orientation is th within 3 degrees."""
    worst = 0.0
    for k in range(12):
        th = k * math.radians(30.0) + 0.1
        img = ob.blob(128, 128, 64.0, 64.0, 6.0)
        y, x = np.mgrid[0:128, 0:128].astype(np.float32)
        t = (x - 64.0) * np.float32(math.cos(th)) + (y - 64.0) * np.float32(math.sin(th))
        img = (img + np.float32(0.02) * t).astype(np.float32)
        g = ob.gaussian_blur(img, 2.0)
        oris = ob.dominant_orientations(g, 64.0, 64.0, 4.0)
        worst = max(worst, min(angle_diff(float(a), th) for a in oris))
    assert math.degrees(worst) < 3.0


def test_descriptor_unit_length():
    """features.rs:1057-1067."""
    kps, desc = ob.detect_and_describe(ob.texture_warped(160, 160, 3, 0.0, 1.0))
    assert len(kps) > 0
    n = np.sqrt((desc.astype(np.float64) ** 2).sum(1))
    assert np.all(np.abs(n - 1.0) < 1e-4)
    assert np.all(desc >= 0.0)


def repeatability(th, sc):
    """features.rs:1013-1039."""
    w = h = 240
    a, _ = ob.detect_and_describe(ob.texture_warped(w, h, 7, 0.0, 1.0), describe=False)
    b, _ = ob.detect_and_describe(ob.texture_warped(w, h, 7, th, sc), describe=False)
    c, s = math.cos(th), math.sin(th)
    hit = n = 0
    for ka in a:
        fx, fy = ka[0] - w / 2.0, ka[1] - h / 2.0
        ex, ey = sc * (c * fx - s * fy) + w / 2.0, sc * (s * fx + c * fy) + h / 2.0
        if ex < 16 or ey < 16 or ex > w - 16 or ey > h - 16:
            continue
        n += 1
        if np.any(np.hypot(b[:, 0] - ex, b[:, 1] - ey) < 2.0):
            hit += 1
    return hit / max(n, 1), n


@pytest.mark.parametrize("deg,sc,min_rate", [(0.0, 1.0, 0.99), (30.0, 1.0, 0.85),
                                             (90.0, 1.0, 0.95), (45.0, 0.8, 0.70),
                                             (0.0, 1.25, 0.70)])
def test_repeatability_under_rotation_and_scale(deg, sc, min_rate):
    """features.rs:1041-1055."""
    rate, n = repeatability(math.radians(deg), sc)
    assert n >= 30
    assert rate >= min_rate


def match_accuracy(th, sc):
    """features.rs:968-1010 (upscale off: the descriptor's own invariance)."""
    w = h = 240
    cfg = ob.detector_cfg(upscale_below=0)
    ka, da = ob.detect_and_describe(ob.texture_warped(w, h, 7, 0.0, 1.0), cfg)
    kb, db = ob.detect_and_describe(ob.texture_warped(w, h, 7, th, sc), cfg)
    c, s = math.cos(th), math.sin(th)
    d2 = ((da[:, None, :].astype(np.float64) - db[None, :, :]) ** 2).sum(-1)
    good = n = 0
    for i in range(len(ka)):
        order = np.argsort(d2[i], kind="stable")
        best, second = d2[i, order[0]], d2[i, order[1]]
        if best >= 0.8 * 0.8 * second:
            continue
        n += 1
        fx, fy = ka[i, 0] - w / 2.0, ka[i, 1] - h / 2.0
        ex, ey = sc * (c * fx - s * fy) + w / 2.0, sc * (s * fx + c * fy) + h / 2.0
        if math.hypot(kb[order[0], 0] - ex, kb[order[0], 1] - ey) < 2.0:
            good += 1
    return good / max(n, 1), n


@pytest.mark.parametrize("deg,sc", [(0.0, 1.0), (30.0, 1.0), (90.0, 1.0), (45.0, 0.8),
                                    (0.0, 1.25)])
def test_matching_under_rotation_and_scale(deg, sc):
    """features.rs:1069-1085."""
    acc, n = match_accuracy(math.radians(deg), sc)
    assert n >= 40
    assert acc >= 0.95


def test_flat_image_has_no_features():
    """features.rs:1211-1219 and :1221-1235."""
    kps, _ = ob.detect_and_describe(np.full((64, 64), 0.5, np.float32))
    assert len(kps) == 0


def test_small_image_upscaled_keeps_original_coords():
    """features.rs:1237-1279."""
    w, h = 240, 160
    centers = [(20.3 + 40.0 * i, 20.7 + 40.0 * j) for i in range(6) for j in range(4)]
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    img = np.full((h, w), 0.2, np.float32)
    for cx, cy in centers:
        r2 = (x - np.float32(cx)) ** 2 + (y - np.float32(cy)) ** 2
        img = (img + np.float32(0.5) * np.exp(-r2 / np.float32(18.0))).astype(np.float32)
    a, _ = ob.detect_and_describe(img, ob.detector_cfg(upscale_below=0))
    b, _ = ob.detect_and_describe(img)
    assert len(b) >= len(centers)
    assert np.all((b[:, 0] >= 0) & (b[:, 0] < w) & (b[:, 1] >= 0) & (b[:, 1] < h))
    near = sum(1 for cx, cy in centers if np.any(np.hypot(b[:, 0] - cx, b[:, 1] - cy) <= 1.0))
    assert near == len(centers)


def test_small_image_upscaled_detects_more():
    """features.rs:1281-1294, on a 320 x 240 textured image (the reference
    renders its synthetic drone scene, which is outside this path)."""
    img = ob.scene_texture(320, 240, 5)
    a, _ = ob.detect_and_describe(img, ob.detector_cfg(upscale_below=0), describe=False)
    b, _ = ob.detect_and_describe(img, describe=False)
    assert len(b) > len(a)


def test_upscale_fill_adds_features_only_when_requested():
    """features.rs:1296-1333, on the same textured image."""
    img = ob.scene_texture(320, 240, 5)
    a, _ = ob.detect_and_describe(img, ob.detector_cfg(upscale_below=0), describe=False)
    cap = len(a) * 3 // 2
    d, _ = ob.detect_and_describe(img, ob.detector_cfg(max_features=cap), describe=False)
    f, _ = ob.detect_and_describe(img, ob.detector_cfg(max_features=cap, upscale_fill=1),
                                  describe=False)
    assert len(d) == len(a)
    assert len(a) < len(f) <= cap
    np.testing.assert_array_equal(f[:len(a), :2], a[:, :2])
    big, _ = ob.detect_and_describe(
        img, ob.detector_cfg(max_features=cap, upscale_fill=1, upscale_below=300), describe=False)
    assert len(big) == min(len(a), cap)


def test_truncation_keeps_the_strongest_in_order():
    """features.rs:777-780: |response| descending, ties in emission order, then
    the first max_features."""
    img = ob.scene_texture(400, 300, 9)
    cfg = ob.detector_cfg(upscale_below=0)
    full, fd = ob.detect_and_describe(img, cfg)
    r = np.abs(full[:, 3])
    assert np.all(np.diff(r) <= 0)
    for m in (1, 7, len(full) // 2):
        part, pd = ob.detect_and_describe(img, ob.detector_cfg(upscale_below=0, max_features=m))
        np.testing.assert_array_equal(part, full[:m])
        np.testing.assert_array_equal(pd, fd[:m])


def test_detector_descriptor_equals_public_describe():
    """The detector's descriptor (features.rs:767) is describe() (:509-519) on
    the keypoint's blurred layer: checked on the first octave, where that
    layer is blur(blur(img, sqrt(1.6^2 - 0.25)), increments) at the keypoint's
    own pixel units."""
    img = ob.scene_texture(200, 160, 3)
    cfg = ob.detector_cfg(upscale_below=0, octaves=1)
    kps, desc = ob.detect_and_describe(img, cfg)
    assert len(kps) > 10
    k = f32(2.0) ** (f32(1.0) / f32(3.0))
    g = ob.gaussian_blur(img, math.sqrt(f32(1.6) * f32(1.6) - f32(0.25)))
    layers = [g]
    for i in range(1, 4):
        prev = f32(1.6) * k ** (i - 1)
        cur = prev * k
        layers.append(ob.gaussian_blur(layers[-1], math.sqrt(cur * cur - prev * prev)))
    checked = 0
    for kp, d in zip(kps, desc):
        x, y, sigma, _, angle = kp
        s = math.log(float(sigma) / 1.6, float(k))  # continuous layer coordinate
        layer = int(round(s))
        if abs(s - layer) > 0.499:  # the refined layer is ambiguous from sigma alone
            continue
        want = ob.describe(layers[layer], x, y, sigma, angle)
        np.testing.assert_array_equal(d, want)
        checked += 1
    assert checked > 10
