"""Parity of skh_build_tracks (the device displacement filter and the host
graph stages) with the oracle: identical tracks (every observation, in
order) and identical TrackStats, bit for bit, for both conflict policies, on
the reference's known-answer cases (tracks.rs:1347-1447) and on synthetic
scenes with dropped and wrong matches."""
import ctypes

import numpy as np
import pytest

import oracle_bind as ob
import workloads

pytestmark = pytest.mark.gpu

import skylens_hip


def _say(msg):
    print(msg, flush=True)


def _cfgs(policy, **kw):
    g = skylens_hip.TrackConfig.default(
        policy=skylens_hip.TRACKS_SPLIT if policy == "split" else skylens_hip.TRACKS_DROP, **kw)
    return g, ob.track_cfg(policy, **kw)


def _equal(ctx, name, pairs, keypoints, policy, **kw):
    _say(f"tracks {name} {policy} {kw}")
    g, o = _cfgs(policy, **kw)
    gt, gs = ctx.build_tracks(pairs, keypoints, g)
    ot, os_ = ob.build_tracks(pairs, keypoints, o)
    assert gs == os_, (name, policy, gs, os_)
    assert gt == ot, (name, policy)
    return gt, gs


@pytest.mark.parametrize("policy", ["drop", "split"])
def test_known_answer_cases(ctx, policy):
    kp = [[(float(i), float(f)) for f in range(3)] for i in range(4)]
    pairs = [(0, 1, [(0, 0), (2, 2)]), (1, 2, [(0, 0), (2, 2)]), (2, 0, [(0, 1)]),
             (2, 3, [(2, 2)])]
    for ml in (2, 3):
        _equal(ctx, "conflicting_component", pairs, kp, policy, min_length=ml)
    kp = [[(0.0, 0.0)] * 2 for _ in range(3)]
    pairs = [(0, 0, [(0, 1)]), (0, 1, [(5, 0), (0, 0)])]
    for ml in (2, 3):
        _equal(ctx, "short_and_invalid", pairs, kp, policy, min_length=ml)


SCENES = [
    ("clean", dict(seed=2, images=10, points=1500, outlier_per_mille=0)),
    ("outliers", dict(seed=3, images=12, points=2500, outlier_per_mille=10)),
    ("sparse_recall", dict(seed=5, images=12, points=2500, keep=0.4, outlier_per_mille=20)),
    ("dense_wrong", dict(seed=8, images=8, points=1200, keep=0.8, outlier_per_mille=60,
                         temporal=3)),
]


@pytest.mark.parametrize("name,kw", SCENES, ids=[s[0] for s in SCENES])
@pytest.mark.parametrize("policy", ["drop", "split"])
def test_scene_equals_oracle(ctx, name, kw, policy):
    s = workloads.track_scene(**kw)
    pairs = [(a, b, m) for a, b, m in s.pairs]
    t, st = _equal(ctx, name, pairs, s.keypoints, policy)
    assert st["tracks"] > 0
    if policy == "split" and kw["outlier_per_mille"]:
        assert st["inconsistent"] > 0  # the device displacement filter flagged matches
    # the cap and a looser length bound
    _equal(ctx, name + "_cap", pairs, s.keypoints, policy, max_tracks=max(1, st["tracks"] // 3))
    _equal(ctx, name + "_len2", pairs, s.keypoints, policy, min_length=2)


def test_empty_and_degenerate_inputs(ctx):
    _equal(ctx, "no_pairs", [], [np.zeros((3, 2))], "split")
    _equal(ctx, "no_keypoints", [(0, 1, [(0, 0)])], [np.zeros((0, 2)), np.zeros((0, 2))], "split")
    # fewer than 8 matches per pair: the displacement filter leaves them alone
    kp = [np.array([[10.0 * k, 5.0 * k] for k in range(6)]) for _ in range(3)]
    pairs = [(0, 1, [(k, k) for k in range(6)]), (1, 2, [(k, (k + 1) % 6) for k in range(6)])]
    _equal(ctx, "seven_or_fewer", pairs, kp, "split", min_length=2)


def test_capacity_and_argument_errors(ctx):
    s = workloads.track_scene(seed=4, images=6, points=600)
    pairs = [(a, b, m) for a, b, m in s.pairs]
    t, st = ctx.build_tracks(pairs, s.keypoints)
    lib, h = ctx.lib, ctx.handle
    ia = np.ascontiguousarray([p[0] for p in pairs], dtype=np.uint64)
    ib = np.ascontiguousarray([p[1] for p in pairs], dtype=np.uint64)
    m_off = np.zeros(len(pairs) + 1, np.uint64)
    m_off[1:] = np.cumsum([len(p[2]) for p in pairs])
    m = np.ascontiguousarray(np.concatenate([p[2] for p in pairs]), dtype=np.uint64)
    kp_off = np.zeros(len(s.keypoints) + 1, np.uint64)
    kp_off[1:] = np.cumsum([len(k) for k in s.keypoints])
    kp = np.ascontiguousarray(np.concatenate(s.keypoints), dtype=np.float64)
    cfg = skylens_hip.TrackConfig.default()
    toff = np.zeros(4, np.uint64)
    obs = np.zeros((8, 2), np.uint64)
    n = ctypes.c_size_t(0)
    stc = skylens_hip.TrackStats()
    args = [h, len(pairs), ia.ctypes.data, ib.ctypes.data, m_off.ctypes.data, m.ctypes.data,
            len(s.keypoints), kp_off.ctypes.data, kp.ctypes.data, ctypes.byref(cfg),
            toff.ctypes.data, 4, obs.ctypes.data, 8, ctypes.byref(n), ctypes.byref(stc)]
    assert lib.skh_build_tracks(*args) == skylens_hip.SKH_EINVAL
    assert n.value == len(t) and stc.tracks == st["tracks"]
    bad = skylens_hip.TrackConfig.default(policy=7)
    args[9] = ctypes.byref(bad)
    assert lib.skh_build_tracks(*args) == skylens_hip.SKH_EINVAL
    # the context stays usable
    t2, st2 = ctx.build_tracks(pairs, s.keypoints)
    assert t2 == t and st2 == st


@pytest.mark.parametrize("positions", [10, 80], ids=["10_positions", "full_scene"])
def test_plan_verified_matches_feed_tracks(ctx, positions):
    """The plan's verified lists (matches kept by RANSAC, tracks.rs:54-69) and
    the batch's keypoints feed build_tracks directly, as the production loop
    does (pipeline.rs:504-508); device result = oracle result. At 80 positions
    this is the full scene (3663 pairs, about 349k verified matches)."""
    import torch

    sc = workloads.full_scene(seed=1, positions=positions)
    desc = torch.from_numpy(sc.desc).cuda()
    kp = torch.from_numpy(np.ascontiguousarray(sc.kp)).cuda()
    plan = skylens_hip.Plan(ctx, sc.offsets, desc.data_ptr(), kp.data_ptr(), sc.pairs, 0.8, True,
                            "bf16")
    try:
        plan.run()
        vl, vij = plan.verified()
        base = plan.match_base
    finally:
        plan.close()
    kps = [sc.view_kp(v).astype(np.float64) for v in range(len(sc.offsets) - 1)]
    pairs = []
    for p, (a, b) in enumerate(sc.pairs):
        m = vij[int(base[p]):int(base[p]) + int(vl[p])].astype(np.int64)
        if len(m):
            pairs.append((int(a), int(b), m))
    assert len(pairs) > 0
    _say(f"plan_verified: {len(pairs)} pairs, {sum(len(p[2]) for p in pairs)} matches")
    for policy in ("split", "drop"):
        t, st = _equal(ctx, "plan_verified", pairs, kps, policy)
        assert st["tracks"] > 0
