"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle on the reference's own test
generators) against the oracle (CPU) and the HIP path (GPU).

The GPU RANSAC cases read inputs and expected outputs from the fixture files
alone; the matching cases regenerate their descriptors with the restated
fixture generator (matching.rs:1580-1656) and compare with the stored lists.
"""
import os

import numpy as np
import pytest

import oracle_bind as ob

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _match_cases():
    z = _load("matches.npz")
    k = 0
    while f"case{k}" in z:
        seed, na, extra, ratio, mutual = z[f"case{k}_args"]
        yield int(seed), int(na), int(extra), float(ratio), bool(mutual), z[f"case{k}"]
        k += 1


def _ransac_cases():
    z = _load("ransac.npz")
    for key in sorted(z):
        if key.endswith("_x1"):
            label = key[:-3]
            yield (label, z[f"{label}_x1"], z[f"{label}_x2"], bool(z[f"{label}_found"][0]),
                   int(z[f"{label}_iters"][0]), z.get(f"{label}_F"), z.get(f"{label}_inliers"))


def test_fixture_files_present():
    assert len(list(_match_cases())) >= 4
    assert len(list(_ransac_cases())) >= 6


@pytest.mark.parametrize("case", list(_match_cases()), ids=lambda c: f"seed{c[0]}_{c[1]}")
def test_oracle_reproduces_match_fixtures(case):
    seed, na, extra, ratio, mutual, want = case
    a, b = ob.match_fixture(seed, na, extra)
    got = np.asarray(ob.ratio_match(a, b, np.float32(ratio), mutual), np.uint32).reshape(-1, 2)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("case", list(_ransac_cases()), ids=lambda c: c[0])
def test_oracle_reproduces_ransac_fixtures(case):
    label, x1, x2, found, iters, F, inl = case
    res = ob.ransac_fundamental(x1, x2)
    assert (res is not None) == found
    assert ob.ransac_last_stats()["iters"] == iters
    if found:
        assert np.array_equal(np.ascontiguousarray(res[0].reshape(9)).view(np.uint64), F)
        assert np.array_equal(res[1], inl)


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(_match_cases()), ids=lambda c: f"seed{c[0]}_{c[1]}")
def test_gpu_reproduces_match_fixtures(ctx, case):
    import skylens_hip

    seed, na, extra, ratio, mutual, want = case
    a, b = ob.match_fixture(seed, na, extra)
    for prec in skylens_hip.available_precisions():
        print(f"golden match seed={seed} na={na} {prec}", flush=True)
        got = np.asarray(ctx.ratio_match(a, b, np.float32(ratio), mutual, prec),
                         np.uint32).reshape(-1, 2)
        assert np.array_equal(got, want), prec


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(_ransac_cases()), ids=lambda c: c[0])
def test_gpu_reproduces_ransac_fixtures(ctx, case):
    import skylens_hip

    label, x1, x2, found, iters, F, inl = case
    print(f"golden ransac {label} n={len(x1)}", flush=True)
    got = ctx.ransac_fundamental(x1, x2, skylens_hip.RansacConfig.default())
    assert (got is not None) == found
    assert ctx.last_ransac_iters == iters
    if found:
        assert np.array_equal(np.ascontiguousarray(got.f.reshape(9)).view(np.uint64), F)
        assert np.array_equal(got.inliers, inl)
