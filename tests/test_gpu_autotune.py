"""The tile autotuner: a tuned table per low precision, plans that mix tiles by
size class (one ranking launch per tile, in sequence on the plan's stream) and
still return the oracle's results for every pair of the full scene and the
uneven batch (fixtures written by the oracle)."""
import numpy as np
import pytest

import scene_golden as sg
import skylens_hip

pytestmark = pytest.mark.gpu

LOWP = [p for p in skylens_hip.available_precisions() if p != "fp32"]


@pytest.mark.parametrize("prec", LOWP)
def test_autotune_table_and_mixed_tile_plans(prec):
    import torch

    ctx = skylens_hip.Context(0)
    try:
        table, ms = ctx.autotune(prec)
        print(f"{prec} tuned table {table.tolist()}", flush=True)
        assert table.shape == (25,) and np.all((table >= 0) & (table < 4))
        assert np.all(ms > 0) and np.all(np.isfinite(ms))
        assert np.array_equal(np.argmin(ms, axis=1), table)
        got, tuned = ctx.tile_table(prec)
        assert tuned and np.array_equal(got, table)
        # force a mixed table (both row-block sizes, both chunk widths) so the
        # grouped launches are exercised whatever the timings chose
        mixed = np.array([(c % 4) for c in range(25)], np.int32)
        for tab in (table, mixed):
            ctx.set_tile_table(prec, tab)
            for name in ("full", "uneven"):
                sc = sg.scene(name)
                z = sg.load(f"scene_{name}.npz")
                desc = torch.from_numpy(sc.desc).cuda()
                kp = torch.from_numpy(np.ascontiguousarray(sc.kp)).cuda()
                plan = skylens_hip.Plan(ctx, sc.offsets, desc.data_ptr(), kp.data_ptr(), sc.pairs,
                                        0.8, True, prec, True, 0.0)
                try:
                    tiles = plan.tiles()
                    print(f"{prec} {name} tiles {tiles}", flush=True)
                    assert sum(tiles.values()) == len(sc.pairs)
                    plan.run()
                    res = plan.results()
                    iters, _, _ = plan.ransac_profile()
                    bad = sg.plan_mismatches(res, z, iters)
                    assert not bad, f"{name} {prec}: {len(bad)} pairs differ, first {bad[:5]}"
                finally:
                    plan.close()
        # a cleared table falls back to the fixed default
        ctx.set_tile_table(prec, None)
        assert not ctx.tile_table(prec)[1]
    finally:
        ctx.close()
