"""The low-precision ranking's error bound against float64 distances.

For the full scene and the uneven batch, at every low precision (BF16, FP8):
every checked key is within the pair's eps of the float64 squared distance it
stands for (the row minimum and its column, the row's second smallest, each
column's two smallest), so every certain decision the bound makes is the exact
one. The largest error and the share of rows sent to the exact re-score are
printed (the measured figures the bench line also reports)."""
import numpy as np
import pytest

import skylens_hip
from fp64_check import scene_errors

pytestmark = pytest.mark.gpu

LOWP = [p for p in skylens_hip.available_precisions() if p != "fp32"]


@pytest.mark.parametrize("prec", LOWP)
@pytest.mark.parametrize("name", ["full", "uneven"])
def test_error_bound_holds(ctx, prec, name):
    import torch

    import scene_golden as sg

    sc = sg.scene(name)
    desc = torch.from_numpy(sc.desc).cuda()
    kp = torch.from_numpy(np.ascontiguousarray(sc.kp)).cuda()
    plan = skylens_hip.Plan(ctx, sc.offsets, desc.data_ptr(), kp.data_ptr(), sc.pairs, 0.8, True,
                            prec, False)
    try:
        plan.run()
        plan.sync()
        st = plan.lowp_state()
    finally:
        plan.close()
    rng = np.random.default_rng(7)
    P = len(sc.pairs)
    # the largest pairs (the bound is loosest there) and a random sample
    size = np.array([(sc.offsets[i + 1] - sc.offsets[i]) * (sc.offsets[j + 1] - sc.offsets[j])
                     for i, j in sc.pairs])
    pick = np.unique(np.concatenate([np.argsort(size)[-6:], rng.choice(P, 14, replace=False)]))
    rep = scene_errors(sc, st, pick)
    rows = int(np.sum(sc.offsets[sc.pairs[:, 0] + 1] - sc.offsets[sc.pairs[:, 0]]))
    print(f"{name} {prec}: max |2key - d2_fp64| = {rep['max_d2_err_fp64']:.3e}, "
          f"max err/eps = {rep['max_err_over_eps']:.3f}, eps median {rep['eps_median']:.3e} "
          f"max {rep['eps_max']:.3e}, re-scored rows {st['queued']}/{rows} "
          f"({st['queued'] / rows:.4f})", flush=True)
    assert rep["max_err_over_eps"] <= 1.0, rep
    assert rep["j1_within_eps"], rep
