"""Pins the CPU oracle's matching half against the reference's own tests.

Each test restates one test of crates/core/src/matching.rs (cited) and runs
it against the oracle, so the oracle is checked before it checks the GPU.
"""
import numpy as np
import pytest

import oracle_bind as ob


# matching.rs:1880-1900 (golden constants :1899-1900)
def test_ratio_match_regression_hash():
    a, b = ob.match_fixture(21, 2000, 700)
    m = ob.ratio_match(a, b, 0.8, True)
    n = ob.ratio_match(a, b, 0.95, False)
    assert (len(m), ob.pairs_hash(m)) == (1145, 0x22CA99B86722152D)
    assert (len(n), ob.pairs_hash(n)) == (1282, 0x7D0CD9993D0C89E5)


# matching.rs:1658-1684: blocked == brute force, both directions, empty inputs.
@pytest.mark.parametrize("seed,na,extra", [(1, 1, 0), (2, 63, 5), (3, 64, 64), (4, 65, 10),
                                           (5, 700, 300)])
def test_ratio_match_equals_reference(seed, na, extra):
    a, b = ob.match_fixture(seed, na, extra)
    for mutual in (True, False):
        for ratio in (0.6, 0.8, 1.0):
            assert ob.ratio_match(a, b, ratio, mutual) == ob.ratio_match_brute(a, b, ratio, mutual)
            assert ob.ratio_match(b, a, ratio, mutual) == ob.ratio_match_brute(b, a, ratio, mutual)


def test_ratio_match_empty():
    a, _ = ob.match_fixture(9, 10, 0)
    empty = np.zeros((0, 128), np.float32)
    assert ob.ratio_match(a, empty, 0.8, True) == []
    assert ob.ratio_match(empty, a, 0.8, True) == []


def _boundary_cases():
    """matching.rs:2359-2405: rows whose f32 ratio decision flips between the
    spec sum order and a sequential sum. Returns (a, b1, b2, r, expect)."""
    g = ob.Lcg(93)
    cases = []
    for _ in range(200):
        a, b1, b2 = g.feature(), g.feature(), g.feature()
        d1 = ob.sq_dist(a, b1, "spec")
        d2 = ob.sq_dist(a, b2, "spec")
        s1 = ob.sq_dist(a, b1, "seq")
        assert ob.sq_dist(a, b1, "impl").view(np.uint32) == d1.view(np.uint32)
        if s1 == d1 or d1 >= d2:
            continue
        target = 0.5 * (float(d1) + float(s1)) / float(d2)
        r = np.float32(np.sqrt(target))

        def between(r):
            lim = np.float32(np.float32(r * r) * d2)
            return (d1 < lim) != (s1 < lim)

        for _ in range(64):
            if between(r):
                break
            r = np.uint32(r.view(np.uint32) + 1).view(np.float32)
        if not between(r):
            continue
        expect = bool(d1 < np.float32(np.float32(r * r) * d2))
        cases.append((a, b1, b2, r, expect))
    return cases


def test_ratio_match_rounding_boundary():
    cases = _boundary_cases()
    assert len(cases) >= 5
    for a, b1, b2, r, expect in cases:
        m = ob.ratio_match(a[None], np.stack([b1, b2]), float(r), False)
        assert (len(m) > 0) == expect


# matching.rs:2407-2425
def test_ratio_match_ties_take_lower_index():
    g = ob.Lcg(11)
    b = np.stack([g.feature() for _ in range(20)])
    for dup in (1, 64 + 6):
        a = np.stack([g.feature() for _ in range(dup + 4)])
        x = b[3].copy()
        x[0] = np.float32(x[0] + np.float32(1e-3))
        a[0] = x
        a[dup] = x
        m = ob.ratio_match(a, b, 0.8, True)
        assert (0, 3) in m
        assert not any(i == dup for i, _ in m)


def _noisy(g, f, scale):
    h = f.copy()
    for k in range(128):
        # Rust `v += 0.01 * g.gauss() as f32` evaluates in f32.
        h[k] = np.float32(h[k] + np.float32(np.float32(scale) * np.float32(_gauss(g))))
    return h


def _gauss(g):
    u1 = max(g.next(), 1e-300)
    u2 = g.next()
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def permutation_fixture():
    """matching.rs:1710-1745 data: a (100), b with 100 noisy copies, 10 twins,
    100 unrelated, deterministically shuffled; returns (a, bf, owner)."""
    g = ob.Lcg(11)
    a = [g.feature() for _ in range(100)]
    b = []
    for i, f in enumerate(a):
        b.append((_noisy(g, f, 0.01), i))
        if i < 10:
            b.append((_noisy(g, f, 0.01), None))
    for _ in range(100):
        b.append((g.feature(), None))
    for i in range(len(b) - 1, 0, -1):
        j = int(g.next() * (i + 1))
        b[i], b[j] = b[j], b[i]
    return np.stack(a), np.stack([x[0] for x in b]), [x[1] for x in b]


def test_ratio_match_recovers_permutation():
    a, bf, owner = permutation_fixture()
    m = ob.ratio_match(a, bf, 0.8, True)
    correct = sum(1 for i, j in m if owner[j] == i)
    assert correct == len(m)
    assert len(m) == 90
    assert all(i >= 10 for i, _ in m)


# matching.rs:1775-1789
def test_ratio_match_needs_second_neighbour():
    g = ob.Lcg(5)
    a = np.stack([g.feature(), g.feature()])
    b = g.feature()[None]
    for mutual in (False, True):
        for ratio in (0.6, 0.8, 1.0):
            assert ob.ratio_match(a, b, ratio, mutual) == []
    b2 = np.stack([a[0], g.feature()])
    assert ob.ratio_match(a[:1], b2, 0.8, False) == [(0, 0)]


def _grid(npos, ncam=3):
    return [(c, p) for p in range(npos) for c in range(ncam)]


# matching.rs:1402-1451 and tests/perf_structure.rs:22-47
def test_candidate_pairs_counts():
    views = _grid(10)
    pairs = ob.candidate_pairs(views, 2, 1, 0)
    assert len(pairs) == 51 + 84
    assert all(i < j for i, j in pairs)
    assert views[pairs[0][0]] == (0, 0)


def test_candidate_pairs_power_of_two_gaps():
    views = _grid(80)
    pairs = ob.candidate_pairs(views, 5, 4, 16)
    same = set()
    for i, j in pairs:
        (ca, pa), (cb, pb) = views[i], views[j]
        if ca == cb:
            same.add(abs(pa - pb))
        else:
            assert abs(pa - pb) <= 4
    assert sorted(same) == [1, 2, 3, 4, 5, 8, 16]
    assert len(pairs) == 1563 + 2100
    assert len(ob.candidate_pairs(views, 5, 4, 2**64 - 1)) == 1563 + 2100 + 3 * 64
    assert len(ob.candidate_pairs(views, 5, 4, 0)) == 1563 + 2100 - 3 * 136


@pytest.mark.parametrize("npos", [20, 40, 80, 160])
def test_candidate_pairs_grow_linearly_with_positions(npos):
    # tests/perf_structure.rs:22-47: 48P - 177 pairs; degree <= 32, mid-path = 32.
    views = _grid(npos)
    pairs = ob.candidate_pairs(views, 5, 4, 16)
    assert len(pairs) == 48 * npos - 177
    if npos == 80:
        deg = np.zeros(len(views), int)
        for i, j in pairs:
            deg[i] += 1
            deg[j] += 1
        assert deg.max() <= 32
        assert deg[3 * 40] == 32
