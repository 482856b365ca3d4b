"""Parity of the HIP ransac_fundamental with the oracle.

Bar: the same Some/None decision, bit-identical F (row-major doubles) and
inlier mask, the same number of hypotheses drawn (the reference's adaptive
stop), and the same model selection (GRIC values to 1e-12 relative: they
pass through log(), whose last bit may differ between glibc and the device
math library). Inputs follow the reference's RANSAC tests in
crates/core/src/matching.rs.
"""
import numpy as np
import pytest

import oracle_bind as ob
import skylens_hip

pytestmark = pytest.mark.gpu


def _cfg(**kw):
    return skylens_hip.RansacConfig.default(**kw)


def _check(ctx, x1, x2, label, **kw):
    print(f"ransac {label} n={len(x1)}", flush=True)
    got = ctx.ransac_fundamental(x1, x2, _cfg(**kw))
    want = ob.ransac_fundamental(x1, x2, ob.default_cfg(**kw))
    iters_o = ob.ransac_last_stats()["iters"]
    assert (got is None) == (want is None), label
    if len(x1) >= 8 and np.all(np.isfinite(x1)) and np.all(np.isfinite(x2)):
        assert ctx.last_ransac_iters == iters_o, label
    if want is None:
        return None
    F, inl, sel = want
    assert np.array_equal(got.f.view(np.uint64), F.view(np.uint64)), label
    assert np.array_equal(got.inliers, inl), label
    assert (got.selection is not None) == bool(sel.has_selection), label
    if sel.has_selection:
        s = got.selection
        assert s.model == ("Homography" if sel.model else "Fundamental"), label
        assert s.parallax == sel.parallax and s.parallax_needed == sel.parallax_needed
        assert s.sigma_px == sel.sigma_px
        assert abs(s.gric_f - sel.gric_f) <= 1e-12 * abs(sel.gric_f)
        assert abs(s.gric_h - sel.gric_h) <= 1e-12 * abs(sel.gric_h)
    return got


@pytest.mark.parametrize("out,seed", [(0.3, 5), (0.5, 9), (0.0, 3), (0.2, 11), (0.6, 13)])
def test_separates_outliers_bitwise(ctx, out, seed):
    x1, x2, truth, _, _ = ob.correspondences(300, 0.5, out, seed)
    got = _check(ctx, x1, x2, f"corr {out} {seed}")
    if out in (0.3, 0.5):
        tp = np.sum(got.inliers & truth)
        assert tp / got.inliers.sum() >= 0.97 and tp / truth.sum() >= 0.98


@pytest.mark.parametrize("d", [1, 2, 3, 7, 19, 40, 76, 100])
def test_many_seed_cases_bitwise(ctx, d):
    x1, x2, _, _, _ = ob.correspondences(300, 0.5, 0.5, d * 7919)
    for r in (1, 5):
        _check(ctx, x1, x2, f"many {d} {r}", seed=r)


def test_4000_correspondences_bitwise(ctx):
    x1, x2, _, _, _ = ob.correspondences(4000, 0.5, 0.5, 4242)
    _check(ctx, x1, x2, "4000")
    assert abs(ctx.last_ransac_iters - 1710) <= 20


def test_degenerate_inputs(ctx):
    x1, x2, _, _, _ = ob.correspondences(50, 0.0, 0.0, 3)
    assert ctx.ransac_fundamental(x1[:7], x2[:7]) is None
    assert ctx.ransac_fundamental(x1, x2[:40]) is None
    for bad in (np.nan, np.inf):
        y = x2.copy()
        y[17, 0] = bad
        assert ctx.ransac_fundamental(x1, y) is None
    i = np.arange(60, dtype=float)
    l1 = np.stack([100 + 13 * i, 200 + 3 * i], 1)
    l2 = np.stack([100 + 13 * i, 350 - 2 * i], 1)
    _check(ctx, l1, l2, "exact lines")
    for sigma in (0.5, 1.0):
        for seed in (1, 3, 5):
            l1, l2 = ob.noisy_lines(sigma, seed)
            _check(ctx, l1, l2, f"noisy lines {sigma} {seed}")
    _check(ctx, x1, x2, "noise free 50")


def test_random_and_non_overlapping(ctx):
    for seed in range(1, 6):
        g = ob.Lcg(seed)
        pts = np.array([[g.next() * 960.0, g.next() * 540.0] for _ in range(400)])
        _check(ctx, pts[:200], pts[200:], f"random {seed}")
        _check(ctx, pts[:200], pts[200:], f"random loose {seed}", min_inlier_ratio=0.0)
    for seed in range(1, 41):
        g = ob.Lcg(seed * 7919)
        n = 15 + int(g.next() * 26.0)
        pts = np.array([[g.next() * 960.0, g.next() * 540.0] for _ in range(2 * n)])
        _check(ctx, pts[:n], pts[n:], f"non-overlap {seed}")
        _check(ctx, pts[:n], pts[n:], f"non-overlap legacy {seed}", min_inliers=8,
               max_chance_prob=0.0)


def test_small_overlap_and_lines_plus_points(ctx):
    for seed in range(1, 11):
        x1, x2, _, _, _ = ob.correspondences(30, 0.5, 1.0 / 3.0, seed)
        _check(ctx, x1, x2, f"small overlap {seed}")
    for k in (2, 6, 10):
        for seed in (1, 2, 3):
            x1, x2, _, _ = ob.line_plus_general(k, 0.5, seed)
            _check(ctx, x1, x2, f"line+{k} {seed}")


def test_planar_and_buildings(ctx):
    for seed in (1, 2, 3):
        x1, x2 = ob.planar_correspondences(200, 0.5, seed)
        got = _check(ctx, x1, x2, f"planar {seed}")
        assert got.is_planar()
    for bld in (0.05, 0.1):
        for seed in (1, 2):
            x1, x2 = ob.ground_with_buildings(300, bld, 0.5, seed)
            got = _check(ctx, x1, x2, f"buildings {bld} {seed}")
            assert not got.is_planar()


def test_single_pair_config(ctx):
    """The single pair's correspondences (the oracle's matches handed off to pixel
    coordinates): same fit as the oracle, and it keeps the true correspondences."""
    from workloads import single_pair

    sp = single_pair(0)
    m = ob.ratio_match(sp.a, sp.b, 0.8, True)
    x1, x2 = sp.handoff(m)
    got = _check(ctx, x1, x2, f"single pair, {len(m)} matches")
    mm = np.asarray(m)
    truth = (sp.b_source[mm[:, 1]] == mm[:, 0]) & sp.b_inlier[mm[:, 1]]
    assert np.sum(got.inliers & truth) / truth.sum() >= 0.95


def test_config_variants(ctx):
    x1, x2, _, _, _ = ob.correspondences(300, 0.5, 0.4, 21)
    _check(ctx, x1, x2, "threshold 1.0", threshold_px=1.0)
    _check(ctx, x1, x2, "max_iters 100", max_iters=100)
    _check(ctx, x1, x2, "seed 77", seed=77)
    _check(ctx, x1, x2, "no chance gate", max_chance_prob=0.0)
