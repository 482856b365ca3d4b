"""The full-size workloads against the oracle's per-pair fixtures.

CPU: the generators still produce the arrays the fixtures were written from
(a drift of numpy or of the generator fails here, not as a parity failure),
and the oracle reproduces the single pair's record and a sample of each
scene's pairs. GPU: every pair of the full scene (3663) and of the uneven
batch (4000) through the batched plan, at every precision, equals the
fixture -- match list, Some/None, F bits, inlier mask, hypotheses drawn; the
single pair runs through the batched plan and through the single-pair C-ABI
(ratio_match, hand-off, ransac_fundamental)."""
import numpy as np
import pytest

import oracle_bind as ob
import scene_golden as sg


def _single_record(sp, matches, fit, iters):
    z = sg.load("single_pair.npz")
    assert len(matches) == int(z["match_count"][0])
    assert ob.pairs_hash(np.asarray(matches, np.uint64)) == int(z["match_hash"][0])
    assert (fit is not None) == bool(z["found"][0])
    assert iters == int(z["iters"][0])
    F, inl = fit
    assert sg.digest(np.ascontiguousarray(np.asarray(F).reshape(9))) == int(z["f_hash"][0])
    assert sg.digest(np.asarray(inl).astype(np.uint8)) == int(z["mask_hash"][0])


def test_single_pair_inputs_and_oracle():
    import workloads

    sp = workloads.single_pair(0)
    z = sg.load("single_pair.npz")
    assert list(z["input_hash"]) == [sg.digest(sp.a), sg.digest(sp.b), sg.digest(sp.kp_a),
                                     sg.digest(sp.kp_b)], "single_pair generator drifted"
    m = ob.ratio_match(sp.a, sp.b, 0.8, True)
    x1, x2 = sp.handoff(m)
    r = ob.ransac_fundamental(x1, x2)
    _single_record(sp, m, (r[0], r[1]), ob.ransac_last_stats()["iters"])
    # the hand-off is real: the matches the fit keeps are mostly true correspondences
    mm = np.asarray(m)
    truth = (sp.b_source[mm[:, 1]] == mm[:, 0]) & sp.b_inlier[mm[:, 1]]
    assert (r[1] & truth).sum() >= 0.95 * truth.sum()
    assert (r[1] & ~truth).sum() <= 0.05 * r[1].sum()


@pytest.mark.parametrize("name", ["full", "uneven"])
def test_scene_inputs_and_oracle_sample(name):
    sc = sg.scene(name)
    z = sg.load(f"scene_{name}.npz")
    assert np.array_equal(sg.scene_inputs_hash(sc), z["input_hash"]), f"{name} generator drifted"
    assert len(sc.pairs) == len(z["match_count"]) == {"full": 3663, "uneven": 4000}[name]
    for p in np.random.default_rng(0).choice(len(sc.pairs), 6, replace=False):
        i, j = (int(v) for v in sc.pairs[p])
        m = ob.ratio_match(sc.view_desc(i), sc.view_desc(j), 0.8, True)
        assert len(m) == z["match_count"][p]
        assert ob.pairs_hash(m) == int(z["match_hash"][p])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["full", "uneven"])
def test_gpu_scene_every_pair(ctx, name):
    import torch

    import skylens_hip

    sc = sg.scene(name)
    z = sg.load(f"scene_{name}.npz")
    assert np.array_equal(sg.scene_inputs_hash(sc), z["input_hash"]), f"{name} generator drifted"
    desc = torch.from_numpy(sc.desc).cuda()
    kp = torch.from_numpy(np.ascontiguousarray(sc.kp)).cuda()
    for prec in skylens_hip.available_precisions():
        print(f"scene {name} {prec}: {len(sc.pairs)} pairs", flush=True)
        plan = skylens_hip.Plan(ctx, sc.offsets, desc.data_ptr(), kp.data_ptr(), sc.pairs, 0.8,
                                True, prec, True, 0.0)
        for rep in range(2):  # a second replay of the graphs gives the same results
            plan.run()
            res = plan.results()
            iters, _, _ = plan.ransac_profile()
            bad = sg.plan_mismatches(res, z, iters)
            assert not bad, f"{name} {prec} replay {rep}: {len(bad)} pairs differ, first {bad[:5]}"
        plan.close()
    del desc, kp


@pytest.mark.gpu
def test_gpu_single_pair_chain(ctx):
    """configs[0] on the GPU: ratio_match -> hand-off -> ransac_fundamental, through the
    single-pair C-ABI at every precision and through a two-view plan."""
    import torch

    import skylens_hip
    import workloads

    sp = workloads.single_pair(0)
    for prec in skylens_hip.available_precisions():
        m = ctx.ratio_match(sp.a, sp.b, 0.8, True, prec)
        x1, x2 = sp.handoff(m)
        fit = ctx.ransac_fundamental(x1, x2, skylens_hip.RansacConfig.default())
        _single_record(sp, m, (fit.f, fit.inliers), ctx.last_ransac_iters)
        desc = torch.from_numpy(np.concatenate([sp.a, sp.b])).cuda()
        kp = torch.from_numpy(np.concatenate([sp.kp_a, sp.kp_b])).cuda()
        plan = skylens_hip.Plan(ctx, np.array([0, 800, 1600], np.uint64), desc.data_ptr(),
                                kp.data_ptr(), np.array([[0, 1]], np.uint32), 0.8, True, prec)
        plan.run()
        res = plan.results()
        iters, _, _ = plan.ransac_profile()
        assert not sg.plan_mismatches(res, sg.load("single_pair.npz"), iters), prec
        plan.close()
