"""Parity of the HIP detector (skh_detect_and_describe) with the oracle.

Bar: bit-exact. Every keypoint field (x, y, sigma, response, angle) and every
descriptor component equal the oracle's, in the oracle's order, for the
reference's own test images (features.rs:819-1335), textured images of
several sizes, both sides of the upscale rule, truncation at max_features,
the upscale_fill rule and non-default configurations.
"""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

import skylens_hip


def _say(msg):
    print(msg, flush=True)


def _cfg(**kw):
    return skylens_hip.DetectorConfig.default(**kw), ob.detector_cfg(**kw)


def _equal(ctx, name, img, **kw):
    _say(f"detector {name} {img.shape} {kw}")
    gcfg, ocfg = _cfg(**kw)
    gk, gd = ctx.detect_and_describe(img, gcfg)
    ok, od = ob.detect_and_describe(img, ocfg)
    assert gk.shape == ok.shape, (name, gk.shape, ok.shape)
    # bit for bit: compare the float bits (NaN-safe)
    np.testing.assert_array_equal(gk.view(np.uint32), ok.view(np.uint32), err_msg=name)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32), err_msg=name)
    return gk


def _blobs_grid(w, h):
    centers = [(20.3 + 40.0 * i, 20.7 + 40.0 * j) for i in range(6) for j in range(4)]
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    img = np.full((h, w), 0.2, np.float32)
    for cx, cy in centers:
        r2 = (x - np.float32(cx)) ** 2 + (y - np.float32(cy)) ** 2
        img = (img + np.float32(0.5) * np.exp(-r2 / np.float32(18.0))).astype(np.float32)
    return img


CASES = [
    ("blob4", lambda: ob.blob(140, 120, 61.0, 47.0, 4.0), {}),
    ("blob8", lambda: ob.blob(140, 120, 70.0, 66.0, 8.0), {}),
    ("blob_subpixel", lambda: ob.blob(140, 120, 61.3, 47.6, 4.0), {}),
    ("texture240", lambda: ob.texture_warped(240, 240, 7, 0.0, 1.0), {}),
    ("texture240_rot30", lambda: ob.texture_warped(240, 240, 7, 0.5235988, 1.0), {}),
    ("texture240_noup", lambda: ob.texture_warped(240, 240, 7, 0.7853982, 0.8),
     {"upscale_below": 0}),
    ("texture160", lambda: ob.texture_warped(160, 160, 3, 0.0, 1.0), {}),
    ("blob_grid_upscaled", lambda: _blobs_grid(240, 160), {}),
    ("scene320_fill", lambda: ob.scene_texture(320, 240, 5),
     {"max_features": 900, "upscale_fill": 1}),
    ("scene320_cap", lambda: ob.scene_texture(320, 240, 5), {"max_features": 37}),
    ("scene640", lambda: ob.scene_texture(640, 480, 2), {}),
    ("scene1024", lambda: ob.scene_texture(1024, 768, 4), {}),
    ("scene1024_cap500", lambda: ob.scene_texture(1024, 768, 4), {"max_features": 500}),
    ("scene_odd", lambda: ob.scene_texture(333, 197, 8), {}),
    ("scales2_oct5", lambda: ob.scene_texture(500, 420, 6),
     {"scales": 2, "octaves": 5, "upscale_below": 0}),
    ("scales4_sigma2", lambda: ob.scene_texture(500, 420, 6),
     {"scales": 4, "sigma0": 2.0, "contrast": 0.005, "edge_ratio": 6.0}),
    ("tiny_only_upscaled", lambda: ob.scene_texture(14, 12, 1, n_blobs=3), {}),
    ("flat", lambda: np.full((64, 64), 0.5, np.float32), {}),
]


@pytest.mark.parametrize("name,make,kw", CASES, ids=[c[0] for c in CASES])
def test_detector_equals_oracle(ctx, name, make, kw):
    img = make()
    k = _equal(ctx, name, img, **kw)
    if name == "flat":
        assert len(k) == 0
    elif name not in ("tiny_only_upscaled",):
        assert len(k) > 0


def test_detector_repeat_runs_identical(ctx):
    _say("detector repeat")
    img = ob.scene_texture(640, 480, 12)
    k1, d1 = ctx.detect_and_describe(img)
    ctx.detect_and_describe(ob.scene_texture(200, 150, 3))  # different sizes in between
    k2, d2 = ctx.detect_and_describe(img)
    np.testing.assert_array_equal(k1.view(np.uint32), k2.view(np.uint32))
    np.testing.assert_array_equal(d1.view(np.uint32), d2.view(np.uint32))


def test_detect_keypoints_only_equals_full(ctx):
    """detect() (features.rs:487-493) is detect_and_describe without descriptors."""
    _say("detector keypoints only")
    img = ob.scene_texture(400, 300, 9)
    k1, _ = ctx.detect_and_describe(img)
    k2, d2 = ctx.detect_and_describe(img, describe=False)
    assert d2 is None
    np.testing.assert_array_equal(k1.view(np.uint32), k2.view(np.uint32))


def test_detector_dev_entry_equals_host_entry(ctx):
    _say("detector device entry")
    import torch

    img = ob.scene_texture(512, 384, 21)
    cfg = skylens_hip.DetectorConfig.default()
    hk, hd = ctx.detect_and_describe(img, cfg)
    d_img = torch.from_numpy(img).cuda()
    cap = int(cfg.max_features)
    d_kp = torch.zeros((cap, 5), dtype=torch.float32, device="cuda")
    d_desc = torch.zeros((cap, 128), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    n = ctx.detect_and_describe_dev(d_img.data_ptr(), img.shape[1], img.shape[0], cfg,
                                    d_kp.data_ptr(), d_desc.data_ptr(), cap)
    torch.cuda.synchronize()
    assert n == len(hk)
    np.testing.assert_array_equal(d_kp[:n].cpu().numpy().view(np.uint32), hk.view(np.uint32))
    np.testing.assert_array_equal(d_desc[:n].cpu().numpy().view(np.uint32), hd.view(np.uint32))


def test_detector_capacity_and_config_errors(ctx):
    _say("detector errors")
    img = ob.scene_texture(300, 200, 4)
    cfg = skylens_hip.DetectorConfig.default()
    kps = np.zeros((3, 5), np.float32)
    import ctypes
    n = ctypes.c_size_t(0)
    rc = ctx.lib.skh_detect_and_describe(ctx.handle, img.ctypes.data, 300, 200, ctypes.byref(cfg),
                                         kps.ctypes.data, None, 3, ctypes.byref(n))
    assert rc == skylens_hip.SKH_EINVAL and n.value > 3
    with pytest.raises(ValueError):
        ctx.detect_and_describe(img, skylens_hip.DetectorConfig.default(scales=0))
    with pytest.raises(ValueError):
        ctx.detect_and_describe(img, skylens_hip.DetectorConfig.default(sigma0=40.0))
    # the context stays usable
    k, _ = ctx.detect_and_describe(img)
    ok, _ = ob.detect_and_describe(img)
    assert len(k) == len(ok)
