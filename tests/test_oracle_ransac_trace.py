"""Pins the oracle's RANSAC loop (crates/core/src/matching.rs:956-1069) to
independent restatements, iteration by iteration, through the oracle's trace.

What each check compares, and the plausible mistake it would catch:

* sampler -- the sample indices against the reference's LCG restated with
  Python integers (matching.rs:956-979): a shifted or truncated LCG output, a
  wrong modulus or duplicate rule fails;
* minimal-sample F -- against LAPACK: the smallest right singular vector of
  the Hartley-normalised 8 x 9 system, rank-2 projection by SVD, denormalised
  and scaled to unit norm (matching.rs:498-557), up to sign; the Some/None
  verdict against numpy's collinearity and eigenvalue-ratio tests; the raw
  inlier count against numpy's Sampson errors (matching.rs:559-573) of the
  oracle's F: a transposed operand, a wrong normaliser, a dropped term fails;
* local optimisation -- fires on exactly the iterations where cnt >= 8 and
  4 cnt >= the best count so far (matching.rs:992-993); each stage's selection
  size against numpy's Sampson errors at th^2 m^2 for m = 3, 2, 1.5, 1, 1, 1,
  its least-squares F against LAPACK, its count; the LO result is the first
  stage reaching the stage maximum when that beats the raw count (strict >,
  matching.rs:1010-1012), and its mask is numpy's Sampson(F_lo) < th^2;
* best count and the iteration limit -- the best count changes only when an
  LO result beats it, and then the limit equals the closed form
  ceil(ln(1 - conf) / ln(1 - w^8)) clamped to [50, max_iters]
  (matching.rs:795-811) computed by numpy; the loop stops exactly at the limit;
* pool -- simulated in Python with the reference's rules (first entry with
  integer Jaccard 10 both > 7 either is the same valley, replaced only by a
  larger count; else appended; Python's stable sort by count, descending;
  truncated to 5, matching.rs:1016-1028) and compared with the oracle's pool
  (counts and the iteration each entry came from) after every iteration: a
  '>=' for '>', an unstable sort or a wrong Jaccard fails;
* tiny inputs (n = 10, 12) -- the best raw consensus equals the maximum over
  the drawn subsets recomputed by numpy;
* final mask -- numpy's Sampson(F_out) < th^2.

Counts compared with numpy allow a difference only for points whose Sampson
error lies within 1e-9 relative of the threshold, where the two evaluation
orders may round differently. The F bits themselves stay unpinned: the
reference takes them from nalgebra's eigen/SVD routines, which do not run here.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob

TH = 1.5
TH2 = TH * TH
M = [3.0, 2.0, 1.5, 1.0, 1.0, 1.0]


# ---------------------------------------------------------------- numpy restatements

def lcg_samples(seed, n, count):
    """matching.rs:956-979 with Python integers: `count` samples of 8 distinct indices."""
    st = (seed ^ 0x9E3779B97F4A7C15) & (2**64 - 1)
    out = []
    for _ in range(count):
        idx = []
        while len(idx) < 8:
            st = (st * 6364136223846793005 + 1442695040888963407) & (2**64 - 1)
            c = (st >> 33) % n
            if c not in idx:
                idx.append(c)
        out.append(idx)
    return out


def sampson(F, x1, x2):
    """matching.rs:559-573, vectorised."""
    a = np.c_[x1, np.ones(len(x1))]
    b = np.c_[x2, np.ones(len(x2))]
    fa = a @ F.T
    ftb = b @ F
    e = np.sum(b * fa, axis=1)
    den = fa[:, 0] ** 2 + fa[:, 1] ** 2 + ftb[:, 0] ** 2 + ftb[:, 1] ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(den <= 0, np.inf, e * e / den)


def count_close(F, x1, x2, t2, want):
    """True when numpy's count of Sampson < t2 equals `want` up to points within
    rounding distance of the threshold."""
    s = sampson(np.asarray(F).reshape(3, 3), x1, x2)
    sure_in = int(np.sum(s < t2 * (1 - 1e-9)))
    maybe = int(np.sum(s < t2 * (1 + 1e-9)))
    return sure_in <= want <= maybe


def mask_close(F, x1, x2, t2, mask):
    s = sampson(np.asarray(F).reshape(3, 3), x1, x2)
    near = np.abs(s - t2) <= 1e-9 * t2
    return bool(np.all((s < t2)[~near] == np.asarray(mask)[~near]))


def collinear(p):
    """matching.rs:224-247."""
    c = p.mean(axis=0)
    d = p - c
    sxx, sxy, syy = np.mean(d[:, 0] ** 2), np.mean(d[:, 0] * d[:, 1]), np.mean(d[:, 1] ** 2)
    tr = sxx + syy
    disc = math.sqrt((sxx - syy) ** 2 + 4 * sxy * sxy)
    hi, lo = math.sqrt(0.5 * (tr + disc)), math.sqrt(max(0.5 * (tr - disc), 0.0))
    return not (hi > 0.0 and lo > 0.02 * hi)


def normalizer(p):
    """matching.rs:498-505."""
    c = p.mean(axis=0)
    d = np.mean(np.linalg.norm(p - c, axis=1))
    s = math.sqrt(2.0) / d if d > 0 else 1.0
    return np.array([[s, 0, -s * c[0]], [0, s, -s * c[1]], [0, 0, 1.0]])


def fundamental_lapack(p, q):
    """matching.rs:507-557 with LAPACK SVDs: (F or None, eigenvalues of A^T A ascending)."""
    if len(p) < 8 or collinear(p) or collinear(q):
        return None, None
    T1, T2 = normalizer(p), normalizer(q)
    a = np.c_[p, np.ones(len(p))] @ T1.T
    b = np.c_[q, np.ones(len(q))] @ T2.T
    A = np.stack([b[:, 0] * a[:, 0], b[:, 0] * a[:, 1], b[:, 0], b[:, 1] * a[:, 0],
                  b[:, 1] * a[:, 1], b[:, 1], a[:, 0], a[:, 1], np.ones(len(p))], 1)
    ev = np.sort(np.linalg.eigvalsh(A.T @ A))
    _, _, vt = np.linalg.svd(A, full_matrices=True)
    Fn = vt[-1].reshape(3, 3)
    u, sv, wt = np.linalg.svd(Fn)
    Fn = u @ np.diag([sv[0], sv[1], 0.0]) @ wt
    F = T2.T @ Fn @ T1
    return F / np.linalg.norm(F), ev


def same_up_to_sign(F, G, tol):
    F = np.asarray(F).reshape(3, 3)
    return min(np.linalg.norm(F - G), np.linalg.norm(F + G)) <= tol


def adaptive_np(w, conf=0.999, mx=2000):
    """matching.rs:795-811: ceil(ln(1 - conf) / ln(1 - w^8)), clamped to [50, mx]."""
    floor = min(50, mx)
    p = min(max(w, 0.0), 1.0) ** 8
    if p >= 1.0:
        return floor
    need = math.log1p(-min(max(conf, 0.0), 1 - 1e-15)) / math.log1p(-p)
    if not math.isfinite(need) or need >= mx:
        return mx
    return min(max(math.ceil(need), floor), mx)


# ---------------------------------------------------------------- cases

def _cases():
    out = []
    for n, out_frac, seed in [(120, 0.3, 1), (300, 0.5, 2), (60, 0.6, 3), (200, 0.2, 7),
                              (400, 0.4, 11)]:
        x1, x2, _, _, _ = ob.correspondences(n, 0.5, out_frac, seed)
        out.append((f"corr_{n}_{out_frac}_{seed}", x1, x2, 2000))
    x1, x2 = ob.planar_correspondences(150, 0.5, 13)
    out.append(("planar_150", x1, x2, 2000))
    x1, x2, _, _ = ob.line_plus_general(20, 0.5, 17)
    out.append(("line_plus_20", x1, x2, 2000))
    x1, x2, _, _, _ = ob.correspondences(250, 0.5, 0.7, 19)
    out.append(("corr_250_0.7_cap300", x1, x2, 300))
    for n, seed in [(10, 23), (12, 29)]:
        x1, x2, _, _, _ = ob.correspondences(n, 0.3, 0.2, seed)
        out.append((f"tiny_{n}", x1, x2, 2000))
    return out


CASES = _cases()


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def traced(request):
    label, x1, x2, mx = request.param
    cfg = ob.default_cfg(max_iters=mx)
    res, recs, masks = ob.ransac_traced(x1, x2, cfg)
    assert recs, label
    return label, x1, x2, cfg, res, recs, masks


def test_sampler_is_the_reference_lcg(traced):
    label, x1, x2, cfg, res, recs, masks = traced
    want = lcg_samples(cfg.seed, len(x1), len(recs))
    assert [list(r.idx) for r in recs] == want


def test_minimal_sample_solutions(traced):
    label, x1, x2, cfg, res, recs, masks = traced
    checked = 0
    for r in recs:
        idx = list(r.idx)
        F, ev = fundamental_lapack(x1[idx], x2[idx])
        if F is None:
            assert not r.ok, (label, r.it, "collinear sample accepted")
            continue
        ratio = ev[1] / ev[8]
        if ratio > 3e-10:  # clear of the degeneracy threshold band
            assert r.ok, (label, r.it, "well-posed sample rejected")
        elif ratio <= 1e-10:
            assert not r.ok, (label, r.it, "degenerate sample accepted")
        if not r.ok:
            continue
        assert count_close(r.F, x1, x2, TH2, r.raw_cnt), (label, r.it)
        if ratio > 1e-6:
            assert same_up_to_sign(r.F, F, 1e-6), (label, r.it)
            checked += 1
    if not label.startswith(("planar", "line")):  # planar scenes: every sample near-degenerate
        assert checked > len(recs) // 4


def test_local_optimisation(traced):
    label, x1, x2, cfg, res, recs, masks = traced
    n_lo = 0
    for k, r in enumerate(recs):
        trigger = bool(r.ok) and r.raw_cnt >= 8 and 4 * r.raw_cnt >= r.best_before
        assert bool(r.lo) == trigger, (label, r.it)
        if not r.lo:
            assert r.lo_cnt == r.raw_cnt
            continue
        n_lo += 1
        cur = np.array(r.F).reshape(3, 3)
        best, best_F = r.raw_cnt, cur
        for s in range(6):
            sel = sampson(cur, x1, x2) < TH2 * M[s] * M[s]
            assert count_close(cur, x1, x2, TH2 * M[s] * M[s], r.stage_ns[s]), (label, r.it, s)
            if not r.stage_ok[s]:
                assert all(r.stage_ns[t] == 0 and not r.stage_ok[t] for t in range(s + 1, 6))
                break
            G = np.array(r.stage_F[s]).reshape(3, 3)
            if sel.sum() == r.stage_ns[s]:
                L, ev = fundamental_lapack(x1[sel], x2[sel])
                if L is not None and ev[1] / ev[8] > 1e-6 and ev[0] < 1e-3 * ev[1]:
                    assert same_up_to_sign(G, L, 1e-6), (label, r.it, s)
            assert count_close(G, x1, x2, TH2, r.stage_cnt[s]), (label, r.it, s)
            if r.stage_cnt[s] > best:
                best, best_F = r.stage_cnt[s], G
            cur = G
        assert r.lo_cnt == best, (label, r.it)
        assert np.array_equal(np.array(r.F_lo).reshape(3, 3), best_F), (label, r.it)
        assert masks[k].sum() == r.lo_cnt
        assert mask_close(r.F_lo, x1, x2, TH2, masks[k]), (label, r.it)
    assert n_lo > 0


def test_best_count_and_iteration_limit(traced):
    label, x1, x2, cfg, res, recs, masks = traced
    n = len(x1)
    best, limit = 0, cfg.max_iters
    for k, r in enumerate(recs):
        assert r.it == k + 1
        assert r.best_before == best, (label, r.it)
        if r.lo and r.lo_cnt > best:
            best = r.lo_cnt
            limit = adaptive_np(best / n, cfg.confidence, cfg.max_iters)
        assert r.iters_after == limit, (label, r.it, r.iters_after, limit)
        last = k == len(recs) - 1
        assert (r.it >= limit) == last, (label, r.it, limit)
    assert ob.ransac_last_stats()["iters"] == len(recs)


def test_pool_follows_the_reference_rules(traced):
    label, x1, x2, cfg, res, recs, masks = traced
    pool = []  # (count, origin iteration, mask)
    for k, r in enumerate(recs):
        if r.lo:
            m = masks[k]
            same = None
            for j, (c, o, pm) in enumerate(pool):
                both = int(np.sum(pm & m))
                either = int(np.sum(pm | m))
                if 10 * both > 7 * either:
                    same = j
                    break
            if same is None:
                pool.append((r.lo_cnt, r.it, m))
            elif pool[same][0] < r.lo_cnt:
                pool[same] = (r.lo_cnt, r.it, m)
            pool = sorted(pool, key=lambda e: -e[0])[:5]
        assert r.npool == len(pool), (label, r.it)
        assert list(r.pool_cnt[:r.npool]) == [e[0] for e in pool], (label, r.it)
        assert list(r.pool_origin[:r.npool]) == [e[1] for e in pool], (label, r.it)
        assert all(a >= b for a, b in zip(r.pool_cnt[:r.npool], r.pool_cnt[1:r.npool]))


def test_tiny_inputs_best_raw_consensus(traced):
    label, x1, x2, cfg, res, recs, masks = traced
    if not label.startswith("tiny"):
        pytest.skip("brute force only on tiny inputs")
    n = len(x1)
    assert cfg.max_iters >= math.comb(n, 8)
    best = 0
    for r in recs:
        F, ev = fundamental_lapack(x1[list(r.idx)], x2[list(r.idx)])
        if F is not None and r.ok:
            best = max(best, int(np.sum(sampson(np.array(r.F).reshape(3, 3), x1, x2) < TH2)))
    assert best == max(r.raw_cnt for r in recs if r.ok)


def test_final_mask_is_sampson_of_f(traced):
    label, x1, x2, cfg, res, recs, masks = traced
    if res is None:
        return
    F, inl, _ = res
    assert mask_close(F, x1, x2, TH2, inl), label
