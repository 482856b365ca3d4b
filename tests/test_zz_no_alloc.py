"""The allocation-free ratio_match entry point against skh_ratio_match_dev.

Two calls in a row through the C-ABI at every precision, the second on a
smaller pair (the shared scratch keeps the first call's contents), each
equal to the allocating entry point's list; a side above the scratch's
8192-row limit is rejected."""
import numpy as np
import pytest

import skylens_hip

pytestmark = pytest.mark.gpu


def _run(fn, ctx, a, b, prec, mutual=True):
    import torch

    da = torch.from_numpy(a).cuda()
    db = torch.from_numpy(b).cuda()
    out = torch.zeros(2 * max(len(a), 1), dtype=torch.int32, device="cuda")
    n = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    fn(da.data_ptr(), len(a), db.data_ptr(), len(b), 0.8, mutual, prec, out.data_ptr(),
       n.data_ptr())
    torch.cuda.synchronize()  # device-wide: the context's stream included
    k = int(n.cpu().numpy()[0])
    return out.cpu().numpy().view(np.uint32)[:2 * k].reshape(-1, 2).copy()


@pytest.mark.parametrize("prec", skylens_hip.available_precisions())
def test_no_alloc_equals_ratio_match_dev(ctx, prec):
    from workloads import scene_pair_descriptors

    cases = [(4097, 3000, 31), (700, 257, 32), (1, 2, 33), (5, 1, 34)]
    for na, nb, seed in cases:  # each call after a larger one
        a, b = scene_pair_descriptors(seed=seed, na=na, nb=nb, pool=max(8, min(na, nb)))
        print(f"no-alloc {prec} {na}x{nb}", flush=True)
        for mutual in (True, False):
            want = _run(ctx.ratio_match_dev, ctx, a, b, prec, mutual)
            got = _run(ctx.ratio_match_dev_noalloc, ctx, a, b, prec, mutual)
            assert np.array_equal(got, want), (prec, na, nb, mutual)
        if na > 100:
            assert len(want) > 0


def test_no_alloc_rejects_oversized(ctx):
    import torch

    a = torch.zeros((8193, 128), dtype=torch.float32, device="cuda")
    out = torch.zeros(2 * 8193, dtype=torch.int32, device="cuda")
    n = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError, match="at most 8192"):
        ctx.ratio_match_dev_noalloc(a.data_ptr(), 8193, a.data_ptr(), 10, 0.8, True, "bf16",
                                    out.data_ptr(), n.data_ptr())
