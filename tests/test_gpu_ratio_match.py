"""Parity of the HIP ratio_match (every precision) with the oracle.

Bar: bit-exact match lists (and, for the exact kernel, bit-exact row/column
state). Cases follow the reference's own tests in crates/core/src/matching.rs.
"""
import numpy as np
import pytest

import oracle_bind as ob
from test_oracle_matching import _boundary_cases, permutation_fixture

pytestmark = pytest.mark.gpu

import skylens_hip

PRECS = skylens_hip.available_precisions()
LOWP = [p for p in PRECS if p != "fp32"]


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("prec", PRECS)
def test_regression_hash(ctx, prec):
    _say(f"regression_hash {prec}")
    a, b = ob.match_fixture(21, 2000, 700)
    m = ctx.ratio_match(a, b, 0.8, True, prec)
    n = ctx.ratio_match(a, b, 0.95, False, prec)
    assert (len(m), ob.pairs_hash(m)) == (1145, 0x22CA99B86722152D)
    assert (len(n), ob.pairs_hash(n)) == (1282, 0x7D0CD9993D0C89E5)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("seed,na,extra", [(1, 1, 0), (2, 63, 5), (3, 64, 64), (4, 65, 10),
                                           (5, 700, 300)])
def test_equals_reference(ctx, prec, seed, na, extra):
    _say(f"equals_reference {prec} {seed} {na} {extra}")
    a, b = ob.match_fixture(seed, na, extra)
    for mutual in (True, False):
        for ratio in (0.6, 0.8, 1.0):
            assert ctx.ratio_match(a, b, ratio, mutual, prec) == ob.ratio_match(a, b, ratio, mutual)
            assert ctx.ratio_match(b, a, ratio, mutual, prec) == ob.ratio_match(b, a, ratio, mutual)


@pytest.mark.parametrize("prec", PRECS)
def test_empty_and_single(ctx, prec):
    a, _ = ob.match_fixture(9, 10, 0)
    empty = np.zeros((0, 128), np.float32)
    assert ctx.ratio_match(a, empty, 0.8, True, prec) == []
    assert ctx.ratio_match(empty, a, 0.8, True, prec) == []
    assert ctx.ratio_match(empty, empty, 0.8, True, prec) == []
    g = ob.Lcg(5)
    a2 = np.stack([g.feature(), g.feature()])
    b1 = g.feature()[None]
    for mutual in (False, True):
        for ratio in (0.6, 0.8, 1.0):
            assert ctx.ratio_match(a2, b1, ratio, mutual, prec) == []
    b2 = np.stack([a2[0], g.feature()])
    assert ctx.ratio_match(a2[:1], b2, 0.8, False, prec) == [(0, 0)]


@pytest.mark.parametrize("prec", PRECS)
def test_rounding_boundary(ctx, prec):
    cases = _boundary_cases()
    assert len(cases) >= 5
    for a, b1, b2, r, expect in cases:
        m = ctx.ratio_match(a[None], np.stack([b1, b2]), float(r), False, prec)
        assert (len(m) > 0) == expect


@pytest.mark.parametrize("prec", PRECS)
def test_ties_take_lower_index(ctx, prec):
    g = ob.Lcg(11)
    b = np.stack([g.feature() for _ in range(20)])
    for dup in (1, 64 + 6, 300):
        a = np.stack([g.feature() for _ in range(dup + 4)])
        x = b[3].copy()
        x[0] = np.float32(x[0] + np.float32(1e-3))
        a[0] = x
        a[dup] = x
        m = ctx.ratio_match(a, b, 0.8, True, prec)
        assert m == ob.ratio_match(a, b, 0.8, True)
        assert (0, 3) in m
        assert not any(i == dup for i, _ in m)


@pytest.mark.parametrize("prec", PRECS)
def test_recovers_permutation(ctx, prec):
    a, bf, owner = permutation_fixture()
    m = ctx.ratio_match(a, bf, 0.8, True, prec)
    assert m == ob.ratio_match(a, bf, 0.8, True)
    assert len(m) == 90


def test_exact_state_bitwise(ctx):
    """Row (j1, d1, d2) and column winner of the exact kernel, bit for bit."""
    for seed, na, extra in [(3, 64, 64), (5, 700, 300), (21, 300, 700)]:
        a, b = ob.match_fixture(seed, na, extra)
        for x, y in ((a, b), (b, a)):
            j1, d1, d2, ci, cd = ctx.ratio_match_state(x, y)
            oj1, od1, od2, oci, ocd = ob.ratio_match_state(x, y)
            none = np.uint64(0xFFFFFFFFFFFFFFFF)
            oj1 = np.where(oj1 == none, 0xFFFFFFFF, oj1).astype(np.uint32)
            oci = np.where(oci == none, 0xFFFFFFFF, oci).astype(np.uint32)
            assert np.array_equal(j1, oj1)
            assert np.array_equal(d1.view(np.uint32), od1.view(np.uint32))
            assert np.array_equal(d2.view(np.uint32), od2.view(np.uint32))
            assert np.array_equal(ci, oci)
            assert np.array_equal(cd.view(np.uint32), ocd.view(np.uint32))


def test_duplicate_rows_and_columns(ctx):
    """Exact duplicates: equal d1 ties in rows (d2 == d1 rejects) and columns
    (lowest i wins), spread across row blocks and column chunks."""
    g = ob.Lcg(77)
    base = np.stack([g.feature() for _ in range(40)])
    a = np.concatenate([base, base[:10], np.stack([g.feature() for _ in range(100)]), base[5:15]])
    b = np.concatenate([base[::-1], np.stack([g.feature() for _ in range(90)]), base[:3]])
    for prec in PRECS:
        for mutual in (True, False):
            for ratio in (0.8, 1.0):
                assert ctx.ratio_match(a, b, ratio, mutual, prec) == \
                    ob.ratio_match(a, b, ratio, mutual), (prec, mutual, ratio)


def test_non_unit_and_extreme_values(ctx):
    """Non-unit descriptors (the golden fixture is non-unit), zeros, a huge
    component, and a NaN row: NaN distances never update any state."""
    rng = np.random.default_rng(1)
    a = rng.standard_normal((150, 128)).astype(np.float32) * 3
    b = np.concatenate([a[:60] + 0.05 * rng.standard_normal((60, 128)).astype(np.float32),
                        rng.standard_normal((70, 128)).astype(np.float32)])
    a[7] = 0.0
    b[3] = 0.0
    a[11, 5] = 1e18
    a[20, 9] = np.nan
    b[50, 0] = np.nan
    for prec in PRECS:
        for mutual in (True, False):
            got = ctx.ratio_match(a, b, 0.8, mutual, prec)
            assert got == ob.ratio_match(a, b, 0.8, mutual), (prec, mutual)


@pytest.mark.parametrize("prec", PRECS)
def test_scene_sized_pair(ctx, prec):
    """A full-scene-sized pair (M, N in the 1500-4096 range) against the oracle."""
    from workloads import scene_pair_descriptors

    a, b = scene_pair_descriptors(seed=4, na=2100, nb=1700)
    _say(f"scene pair {prec} {a.shape} {b.shape}")
    got = ctx.ratio_match(a, b, 0.8, True, prec)
    want = ob.ratio_match(a, b, 0.8, True)
    assert got == want
    assert len(want) > 50


@pytest.mark.parametrize("prec", LOWP)
@pytest.mark.parametrize("tile", ["128x64", "128x128", "256x64", "256x128"])
def test_every_tile_shape(ctx, prec, tile):
    """Each MFMA tile gives the exact list (ragged row blocks and column chunks)."""
    _say(f"tile {tile} {prec}")
    ctx.set_tile(skylens_hip.TILES[tile])
    try:
        for seed, na, extra in [(3, 64, 64), (5, 700, 300), (21, 2000, 700)]:
            a, b = ob.match_fixture(seed, na, extra)
            for mutual in (True, False):
                assert ctx.ratio_match(a, b, 0.8, mutual, prec) == ob.ratio_match(a, b, 0.8, mutual)
                assert ctx.ratio_match(b, a, 0.8, mutual, prec) == ob.ratio_match(b, a, 0.8, mutual)
        from workloads import scene_pair_descriptors
        a, b = scene_pair_descriptors(seed=11, na=1537, nb=4097)
        assert ctx.ratio_match(a, b, 0.8, True, prec) == ob.ratio_match(a, b, 0.8, True)
        # one partly filled chunk, a chunk whose second column tile is all
        # padding, and a last chunk of one column
        for seed, na, nb in [(12, 17, 5), (13, 33, 95), (14, 300, 31), (15, 129, 193)]:
            a, b = scene_pair_descriptors(seed=seed, na=na, nb=nb, pool=48)
            assert ctx.ratio_match(a, b, 0.8, True, prec) == ob.ratio_match(a, b, 0.8, True)
            assert ctx.ratio_match(b, a, 0.8, False, prec) == ob.ratio_match(b, a, 0.8, False)
    finally:
        ctx.set_tile(-1)


@pytest.mark.parametrize("prec", LOWP)
@pytest.mark.parametrize("tile", ["128x64", "128x128", "256x64", "256x128"])
def test_large_finite_descriptors(ctx, prec, tile):
    """Large but finite components: |a|^2 far above 1 (near and past the range the
    MFMA ranking accepts), nb = 63 so every chunk has padding columns, na = 300 so
    the last row block has padding rows. The padding keys must never enter a
    top-2, in mutual and non-mutual mode."""
    _say(f"large finite {prec} {tile}")
    ctx.set_tile(skylens_hip.TILES[tile])
    try:
        for scale in (2e14, 5e12, 1e6):
            rng = np.random.default_rng(int(np.log10(scale)))
            a = (rng.standard_normal((300, 128)) * scale).astype(np.float32)
            b = np.concatenate([a[:40] + np.float32(0.05 * scale) *
                                rng.standard_normal((40, 128)).astype(np.float32),
                                (rng.standard_normal((23, 128)) * scale).astype(np.float32)])
            for mutual in (True, False):
                for x, y in ((a, b), (b, a)):
                    got = ctx.ratio_match(x, y, 0.8, mutual, prec)
                    assert got == ob.ratio_match(x, y, 0.8, mutual), (scale, mutual, x.shape)
    finally:
        ctx.set_tile(-1)
