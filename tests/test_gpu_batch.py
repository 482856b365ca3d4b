"""Parity of the batched per-pair driver (one captured graph for all pairs)
with the oracle run pair by pair (benches/pipeline.rs:247-295 order):
ratio_match(view i, view j, 0.8, mutual) bit-exact, then ransac_fundamental on
the gathered keypoints (f32 -> f64, + px_offset) with identical F, mask and
Some/None.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_bind as ob
import skylens_hip

pytestmark = pytest.mark.gpu


def _scene(seed, positions, lo, hi, pool=1500):
    from workloads import scene

    rng = np.random.default_rng(seed)
    counts = rng.integers(lo, hi, 3 * positions)
    # a small descriptor pool so that small views still share enough features
    return scene(seed, positions, counts, 0.106, 0.368, pool_size=pool)


def _oracle_pair(sc, p, px_offset, mutual):
    i, j = (int(v) for v in sc.pairs[p])
    a, b = sc.view_desc(i), sc.view_desc(j)
    m = ob.ratio_match(a, b, 0.8, mutual)
    mm = np.array(m, dtype=np.int64).reshape(-1, 2)
    x1 = sc.view_kp(i)[mm[:, 0]].astype(np.float64) + px_offset
    x2 = sc.view_kp(j)[mm[:, 1]].astype(np.float64) + px_offset
    return m, ob.ransac_fundamental(x1, x2)


def _run_plan(ctx, sc, prec, px_offset=0.0, mutual=True, steps=2):
    import torch

    desc = torch.from_numpy(sc.desc).cuda()
    kp = torch.from_numpy(np.ascontiguousarray(sc.kp)).cuda()
    plan = skylens_hip.Plan(ctx, sc.offsets, desc.data_ptr(), kp.data_ptr(), sc.pairs, 0.8,
                            mutual, prec, True, px_offset)
    outs = []
    for _ in range(steps):  # replays must agree with each other
        plan.run()
        plan.sync()
        outs.append(plan.results())
    plan.close()
    del desc, kp
    return outs


@pytest.mark.parametrize("prec", skylens_hip.available_precisions())
def test_batch_matches_oracle(ctx, prec):
    sc = _scene(5, 8, 150, 700)
    print(f"batch {prec}: {len(sc.views)} views, {len(sc.pairs)} pairs", flush=True)
    r1, r2 = _run_plan(ctx, sc, prec)
    with ThreadPoolExecutor(8) as ex:
        want = list(ex.map(lambda p: _oracle_pair(sc, p, 0.0, True), range(len(sc.pairs))))
    n_found = 0
    for p, (m, rf) in enumerate(want):
        got = [tuple(map(int, x)) for x in r1.matches(p)]
        assert got == m, f"pair {p}"
        assert r1.found[p] == (1 if rf is not None else 0), f"pair {p}"
        if rf is not None:
            n_found += 1
            F, inl, sel = rf
            assert np.array_equal(r1.F[p].view(np.uint64), F.view(np.uint64)), f"pair {p}"
            assert np.array_equal(r1.inlier_mask(p), inl), f"pair {p}"
    assert n_found > len(sc.pairs) // 3
    # a second replay of the graph gives the same bytes
    assert np.array_equal(r1.match_len, r2.match_len)
    assert np.array_equal(r1.F.view(np.uint64), r2.F.view(np.uint64))
    assert np.array_equal(r1.found, r2.found)


def test_batch_pixel_offset_and_non_mutual(ctx):
    sc = _scene(7, 6, 100, 400)
    prec = skylens_hip.available_precisions()[-1]
    (r,) = _run_plan(ctx, sc, prec, px_offset=0.5, mutual=False, steps=1)
    for p in range(0, len(sc.pairs), 5):
        m, rf = _oracle_pair(sc, p, 0.5, False)
        assert [tuple(map(int, x)) for x in r.matches(p)] == m
        assert r.found[p] == (1 if rf is not None else 0)
        if rf is not None:
            assert np.array_equal(r.F[p].view(np.uint64), rf[0].view(np.uint64))


def test_batch_empty_and_tiny_views(ctx):
    """Views with 0, 1 and 2 features; pairs whose match count is below 8."""
    import torch

    from workloads import unit_rows

    rng = np.random.default_rng(3)
    counts = [0, 1, 2, 9, 40, 300]
    desc = np.concatenate([unit_rows(rng, c) for c in counts if c] or [np.zeros((0, 128), np.float32)])
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    kp = rng.uniform(0, 500, (int(offsets[-1]), 2)).astype(np.float32)
    pairs = np.array([(i, j) for i in range(6) for j in range(6) if i != j], np.uint32)
    d = torch.from_numpy(desc).cuda()
    k = torch.from_numpy(kp).cuda()
    for prec in skylens_hip.available_precisions():
        plan = skylens_hip.Plan(ctx, offsets, d.data_ptr(), k.data_ptr(), pairs, 0.8, True, prec)
        plan.run()
        r = plan.results()
        for p, (i, j) in enumerate(pairs):
            a = desc[offsets[i]:offsets[i + 1]]
            b = desc[offsets[j]:offsets[j + 1]]
            assert [tuple(map(int, x)) for x in r.matches(p)] == ob.ratio_match(a, b, 0.8, True)
        plan.close()


def test_plans_interleaved(ctx):
    """Four plans on four contexts (four streams), runs issued in rotation with no
    synchronisation in between (the way bench.py issues steps): every run's
    results equal an isolated run's, byte for byte."""
    import torch

    sc = _scene(11, 8, 150, 900)
    prec = skylens_hip.available_precisions()[-1]
    (ref,) = _run_plan(ctx, sc, prec, steps=1)
    desc = torch.from_numpy(sc.desc).cuda()
    kp = torch.from_numpy(np.ascontiguousarray(sc.kp)).cuda()
    extra = [skylens_hip.Context(0) for _ in range(3)]
    plans = [skylens_hip.Plan(c, sc.offsets, desc.data_ptr(), kp.data_ptr(), sc.pairs, 0.8, True,
                              prec) for c in [ctx] + extra]
    for k in range(10):
        plans[k % len(plans)].run()
    P = len(sc.pairs)
    for plan in plans:
        r = plan.results()
        assert np.array_equal(r.match_len, ref.match_len)
        assert all(np.array_equal(r.matches(p), ref.matches(p)) for p in range(P))
        assert np.array_equal(r.found, ref.found)
        for p in range(P):
            if ref.found[p] == 1:
                assert np.array_equal(r.F[p].view(np.uint64), ref.F[p].view(np.uint64)), f"pair {p}"
                assert np.array_equal(r.inlier_mask(p), ref.inlier_mask(p)), f"pair {p}"
        plan.close()
    for c in extra:
        c.close()


def test_plan_verified_matches(ctx):
    """The plan's device-side verified lists (tracks.rs:56-69 per pair) equal each
    pair's matches filtered by its inlier mask, and are empty without a model."""
    import torch

    sc = _scene(13, 8, 150, 900)
    desc = torch.from_numpy(sc.desc).cuda()
    kp = torch.from_numpy(np.ascontiguousarray(sc.kp)).cuda()
    for prec in skylens_hip.available_precisions():
        plan = skylens_hip.Plan(ctx, sc.offsets, desc.data_ptr(), kp.data_ptr(), sc.pairs, 0.8,
                                True, prec)
        try:
            plan.run()
            r = plan.results()
            vl, vij = plan.verified()
            for p in range(len(sc.pairs)):
                b = int(r.match_base[p])
                if r.found[p] == 1:
                    want = skylens_hip.verified_matches(r.matches(p), r.inlier_mask(p))
                else:
                    want = []
                got = [tuple(map(int, x)) for x in vij[b:b + int(vl[p])]]
                assert got == want, (prec, p)
            assert int(vl.sum()) > 0
        finally:
            plan.close()
