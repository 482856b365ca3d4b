"""scheduled_pairs (matching.rs:47-142) and verified_matches (tracks.rs:56-69):
the reference's own tests (matching.rs:2668-2700, sparse.rs:1140-1170,
tracks.rs:1335-1343) on the oracle and the C-ABI, and the C-ABI against the
oracle on random views and schedules."""
import numpy as np
import pytest

import oracle_bind as ob
import skylens_hip

FRONT, RIGHT, LEFT = 0, 1, 2


def _grid(npos):
    return [(c, p) for p in range(npos) for c in range(3)]


def _sched_c(kind="formation", **kw):
    s = skylens_hip.PairSchedule.spec() if kind == "spec" else skylens_hip.PairSchedule.default()
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _connected(n, pairs):
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for i, j in pairs:
        parent[find(i)] = find(j)
    return len({find(x) for x in range(n)}) == 1


@pytest.mark.parametrize("impl", ["oracle", "capi"])
def test_spec_schedule_is_candidate_pairs_and_formation_offsets(impl):
    """matching.rs:2668-2700."""
    views = _grid(30)
    if impl == "oracle":
        spec = ob.scheduled_pairs(views, ob.pair_schedule("spec"))
        f = ob.scheduled_pairs(views, ob.pair_schedule())
        cand = ob.candidate_pairs(views, 5, 4, 16)
    else:
        spec = skylens_hip.scheduled_pairs(views, _sched_c("spec"))
        f = skylens_hip.scheduled_pairs(views)
        cand = skylens_hip.candidate_pairs(views, 5, 4, 16)
    assert spec == cand
    assert all(i < j for i, j in f)
    kinds = set()
    for i, j in f:
        (ca, pa), (cb, pb) = views[i], views[j]
        if ca != cb:
            (_, pf), (co, po) = ((ca, pa), (cb, pb)) if ca == FRONT else ((cb, pb), (ca, pa))
            assert po >= pf + 20 and po <= pf + 40 and (po - pf) % 4 == 0
            kinds.add(co)
    assert sorted(kinds) == [RIGHT, LEFT]
    same = lambda v: sum(1 for i, j in v if views[i][0] == views[j][0])  # noqa: E731
    assert same(f) == same(spec)


@pytest.mark.parametrize("impl", ["oracle", "capi"])
def test_default_schedule_is_connected(impl):
    """sparse.rs:1140-1170."""
    views = _grid(44)
    if impl == "oracle":
        pairs = ob.scheduled_pairs(views, ob.pair_schedule())
        t = ob.scheduled_pairs(views, ob.pair_schedule(right_min=1, left_min=1, max=0, step=1))
    else:
        pairs = skylens_hip.scheduled_pairs(views)
        t = skylens_hip.scheduled_pairs(views, _sched_c(right_min=1, left_min=1, max=0, step=1))
    assert _connected(len(views), pairs)
    has = lambda a, b: any({views[i][0], views[j][0]} == {a, b} for i, j in pairs)  # noqa: E731
    assert has(0, 1) and has(0, 2) and not has(1, 2)
    assert not _connected(len(views), t)


def test_capi_equals_oracle_on_random_views_and_schedules():
    rng = np.random.default_rng(5)
    for trial in range(40):
        n = int(rng.integers(0, 90))
        views = [(int(rng.integers(0, 4)), int(rng.integers(0, 60))) for _ in range(n)]
        kw = dict(temporal=int(rng.integers(0, 7)), pow2_max=int(rng.choice([0, 4, 16, 64])),
                  right_min=int(rng.integers(0, 25)), left_min=int(rng.integers(0, 25)),
                  max=int(rng.integers(0, 45)), step=int(rng.integers(0, 6)),
                  range=int(rng.integers(0, 6)))
        kind = "spec" if trial % 3 == 0 else "formation"
        want = ob.scheduled_pairs(views, ob.pair_schedule(kind, **kw))
        got = skylens_hip.scheduled_pairs(views, _sched_c(kind, **kw))
        assert got == want, (trial, kind, kw)


def test_schedule_counts_on_the_full_scene_grid():
    views = _grid(80)
    assert len(skylens_hip.scheduled_pairs(views, _sched_c("spec"))) == 3663
    f = skylens_hip.scheduled_pairs(views)
    assert f == ob.scheduled_pairs(views, ob.pair_schedule())


@pytest.mark.parametrize("impl", ["oracle", "capi"])
def test_verified_matches(impl):
    """tracks.rs:1335-1343."""
    m = [(0, 1), (2, 3), (4, 5)]
    if impl == "oracle":
        assert ob.verified_matches(m, [True, False, True]) == [(0, 1), (4, 5)]
        assert ob.verified_matches(m, [True, False]) == ("err", (3, 2))
        assert ob.verified_matches(m[:1], [True, True]) == ("err", (1, 2))
    else:
        assert skylens_hip.verified_matches(m, [True, False, True]) == [(0, 1), (4, 5)]
        with pytest.raises(skylens_hip.LengthMismatch) as e:
            skylens_hip.verified_matches(m, [True, False])
        assert e.value.lengths == (3, 2)
        with pytest.raises(skylens_hip.LengthMismatch) as e:
            skylens_hip.verified_matches(m[:1], [True, True])
        assert e.value.lengths == (1, 2)
        assert skylens_hip.verified_matches([], []) == []
