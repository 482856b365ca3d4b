"""Pins the CPU oracle's RANSAC half against the reference's own tests.

Each test restates one test of crates/core/src/matching.rs (cited), run on
the oracle. The exact F bits depend on the eigen/SVD algorithms (nalgebra in
the reference, cyclic Jacobi here), so these tests pin the oracle at the
reference's level: closed forms, known answers, and the statistical suites.
The own solvers are additionally checked against numpy's LAPACK.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_bind as ob

TH2 = 1.5 ** 2


# matching.rs:1389-1400
def test_adaptive_iterations_does_not_collapse():
    assert ob.adaptive_iterations(2.0 / 300.0, 8, 0.999, 2000) == 2000
    assert ob.adaptive_iterations(0.0, 8, 0.999, 2000) == 2000
    assert ob.adaptive_iterations(1.0, 8, 0.999, 2000) == 50
    assert ob.adaptive_iterations(0.5, 8, 0.999, 2000) == 1765
    assert ob.adaptive_iterations(0.95, 8, 0.999, 2000) == 50


def test_adaptive_iterations_closed_form():
    for w in np.linspace(0.05, 0.99, 60):
        for s in (4, 8):
            need = np.log1p(-0.999) / np.log1p(-(w ** s))
            want = 2000 if need >= 2000 else int(min(max(np.ceil(need), 50), 2000))
            assert ob.adaptive_iterations(float(w), s, 0.999, 2000) == want


# matching.rs:2265-2272
def test_binomial_tail_matches_direct_sum():
    assert abs(ob.binomial_tail(5, 2, 0.3) - 0.47178) < 1e-12
    assert ob.binomial_tail(5, 0, 0.3) == 1.0
    assert ob.binomial_tail(5, 6, 0.3) == 0.0
    assert abs(ob.binomial_tail(10, 10, 0.5) - 0.5 ** 10) < 1e-15
    from math import comb
    for n, k, p in [(30, 7, 0.02), (60, 20, 0.3), (12, 1, 0.5)]:
        direct = sum(comb(n, i) * p ** i * (1 - p) ** (n - i) for i in range(k, n + 1))
        assert abs(ob.binomial_tail(n, k, p) - direct) < 1e-12


# matching.rs:2806-2819
def test_parallax_needed_values():
    assert [ob.parallax_needed(n) for n in (50, 200, 300, 1000)] == [4, 6, 6, 10]
    for n in (50, 300, 1000):
        k = ob.parallax_needed(n)
        assert ob.binomial_tail(n, k, 0.001) < 1e-6
        assert ob.binomial_tail(n, k - 1, 0.001) >= 1e-6


def test_eigen_min_against_lapack():
    # eigenvalues against LAPACK; the vector is a unit eigenvector of the
    # smallest one (A v = lambda v), the same line as LAPACK's when the gap
    # to the next eigenvalue is wide
    rng = np.random.default_rng(2)
    for trial in range(60):
        X = rng.standard_normal((9 if trial % 2 else 12, 9)) * rng.uniform(0.1, 10, 9)
        A = X.T @ X
        ev, v = ob.eigen_min9(A)
        ref, V = np.linalg.eigh(A)
        assert np.allclose(np.sort(ev), ref, rtol=1e-10, atol=1e-10 * ref[-1])
        assert abs(np.linalg.norm(v) - 1.0) < 1e-13
        assert np.linalg.norm(A @ v - ref[0] * v) < 1e-9 * ref[-1]
        if ref[1] - ref[0] > 1e-3 * ref[-1]:
            assert min(np.linalg.norm(v - V[:, 0]), np.linalg.norm(v + V[:, 0])) < 1e-6


def test_homography_4pt_closed_form():
    # four correspondences fix H: the closed form returns the generating H (up to
    # sign, unit norm), maps every sample point onto its partner, equals the exact
    # null vector of the 8x9 DLT system (numpy SVD of the unnormalised system, to
    # its own accuracy), agrees with the A^T A
    # eigenvector route within that route's own error (it squares the condition
    # number), and rejects collinear triples
    rng = np.random.default_rng(11)
    for trial in range(200):
        x1 = rng.uniform(0, 1000, (4, 2))
        Ht = np.eye(3) + rng.normal(0, 0.2, (3, 3))
        Ht[2, :2] *= 1e-3
        q = np.c_[x1, np.ones(4)] @ Ht.T
        x2 = q[:, :2] / q[:, 2:]
        H = ob.homography_4pt(x1, x2)
        assert H is not None
        assert abs(np.linalg.norm(H) - 1.0) < 1e-12
        T = Ht / np.linalg.norm(Ht)
        assert min(np.linalg.norm(H - T), np.linalg.norm(H + T)) < 1e-8
        p = np.c_[x1, np.ones(4)] @ H.T
        assert np.max(np.abs(p[:, :2] / p[:, 2:] - x2)) < 1e-8
        G = ob.homography_dlt(x1, x2)
        assert min(np.linalg.norm(H - G), np.linalg.norm(H + G)) < 1e-3
        A = []
        for (x, y), (u, v) in zip(x1, x2):
            A.append([-x, -y, -1, 0, 0, 0, u * x, u * y, u])
            A.append([0, 0, 0, -x, -y, -1, v * x, v * y, v])
        h0 = np.linalg.svd(np.array(A))[2][-1]
        h0 = h0 / np.linalg.norm(h0)
        assert min(np.linalg.norm(H.ravel() - h0), np.linalg.norm(H.ravel() + h0)) < 1e-6
    col = np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 2.0], [5.0, 1.0]])
    assert ob.homography_4pt(col, col + 3.0) is None


def test_eigen_min_inverse_iteration():
    # the 8-point solver's eigen step (inverse iteration, error ~ (l1/l2)^steps):
    # LAPACK's smallest eigenvector (up to sign) when the gap is wide, a small
    # residual when l1 < l2 / 2, a Rayleigh quotient below l2 always; degenerate
    # exactly when two eigenvalues sit below 1e-10 of the largest
    rng = np.random.default_rng(4)
    for trial in range(100):
        X = rng.standard_normal((12, 9)) * rng.uniform(0.1, 10, 9)
        A = X.T @ X
        deg, v = ob.eigen_min_inv9(A)
        assert not deg
        w, V = np.linalg.eigh(A)
        assert abs(np.linalg.norm(v) - 1.0) < 1e-13
        assert v @ A @ v < w[1]
        if w[0] < 0.5 * w[1]:
            assert np.linalg.norm(A @ v - w[0] * v) < 1e-9 * w[-1]
        if w[0] < 0.1 * w[1]:
            assert min(np.linalg.norm(v - V[:, 0]), np.linalg.norm(v + V[:, 0])) < 1e-10
    Q, _ = np.linalg.qr(rng.standard_normal((9, 9)))
    for lam2, want in ((1e-11, True), (1e-9, False), (0.0, True)):
        lam = np.array([0.0, lam2, 0.3, 0.5, 0.7, 0.8, 0.9, 0.95, 1.0])
        deg, v = ob.eigen_min_inv9((Q * lam) @ Q.T)
        assert deg == want, lam2
    lam = np.array([0.0, 1e-3, 0.3, 0.5, 0.7, 0.8, 0.9, 0.95, 1.0])
    deg, v = ob.eigen_min_inv9((Q * lam) @ Q.T)
    assert not deg and min(np.linalg.norm(v - Q[:, 0]), np.linalg.norm(v + Q[:, 0])) < 1e-9


def test_eigen_min_null_vectors():
    # rank-deficient A = X^T X (8 rows): the vector spans the null space of X,
    # also for a tridiagonal form whose eliminations pivot (graded rows) and
    # for a block-diagonal A whose tridiagonal form splits
    rng = np.random.default_rng(5)
    for trial in range(40):
        X = rng.standard_normal((8, 9)) * (10.0 ** rng.uniform(-2, 2, 9))
        A = X.T @ X
        ev, v = ob.eigen_min9(A)
        assert abs(np.linalg.norm(v) - 1.0) < 1e-13
        assert np.linalg.norm(X @ v) < 1e-9 * np.linalg.norm(X)
        assert np.min(ev) < 1e-9 * np.max(ev)
    D = np.diag([5.0, 1.0, 7.0, 2.0, 3.0, 0.5, 9.0, 4.0, 6.0])
    ev, v = ob.eigen_min9(D)
    assert np.allclose(np.sort(ev), np.sort(np.diag(D)), rtol=0, atol=1e-14)
    assert abs(abs(v[5]) - 1.0) < 1e-14 and np.linalg.norm(np.delete(v, 5)) < 1e-14


# matching.rs:1222-1235
def test_eight_point_exact_on_noise_free_points():
    x1, x2, _, c1, c2 = ob.correspondences(50, 0.0, 0.0, 3)
    F = ob.fundamental_8pt(x1, x2)
    worst = max(ob.sampson_error(F, x1[i], x2[i]) ** 0.5 for i in range(50))
    assert worst < 1e-6
    G = ob.fundamental_from_cameras(c1, c2)
    assert min(np.linalg.norm(F - G), np.linalg.norm(F + G)) < 1e-6
    assert abs(np.linalg.det(F)) < 1e-9
    # rank 2 with unit Frobenius norm, as numpy's SVD sees it
    s = np.linalg.svd(F, compute_uv=False)
    assert s[2] < 1e-12 and abs(np.linalg.norm(F) - 1.0) < 1e-12


def test_sampson_error_closed_form():
    rng = np.random.default_rng(4)
    for _ in range(50):
        F = rng.standard_normal((3, 3))
        p, q = rng.uniform(0, 900, 2), rng.uniform(0, 500, 2)
        a, b = np.r_[p, 1.0], np.r_[q, 1.0]
        fa, ftb = F @ a, F.T @ b
        want = (b @ fa) ** 2 / (fa[0] ** 2 + fa[1] ** 2 + ftb[0] ** 2 + ftb[1] ** 2)
        assert abs(ob.sampson_error(F, p, q) - want) <= 1e-12 * max(1.0, want)


# matching.rs:2086-2123
def test_sampson_jacobian_matches_numeric():
    x1, x2, _, c1, c2 = ob.correspondences(60, 0.5, 0.3, 17)
    t1, t2 = ob.normalizer(x1), ob.normalizer(x2)
    s1, s2 = t1[0, 0], t2[0, 0]
    f = ob.fundamental_from_cameras(c1, c2)
    g0 = np.linalg.inv(t2).T @ f @ np.linalg.inv(t1)
    rng = ob.Lcg(5)
    shake = np.array([[_g(rng) * 0.01 * np.linalg.norm(g0) for _ in range(3)] for _ in range(3)])
    worst = 0.0
    for g in (g0 / np.linalg.norm(g0), (g0 + shake) / np.linalg.norm(g0 + shake)):
        for p, q in zip(x1, x2):
            a = t1 @ np.r_[p, 1.0]
            b = t2 @ np.r_[q, 1.0]
            r, j = ob.sampson_residual_jacobian(g, s1, s2, a, b)
            fp = t2.T @ g @ t1
            assert abs(r * r - ob.sampson_error(fp, p, q)) <= 1e-9 * (1 + r * r)
            h = 1e-6
            num = diff = 0.0
            for k in range(9):
                gp, gm = g.copy(), g.copy()
                gp[k // 3, k % 3] += h
                gm[k // 3, k % 3] -= h
                d = (ob.sampson_residual_jacobian(gp, s1, s2, a, b)[0]
                     - ob.sampson_residual_jacobian(gm, s1, s2, a, b)[0]) / (2 * h)
                num += d * d
                diff += (d - j[k]) ** 2
            worst = max(worst, np.sqrt(diff) / np.sqrt(num))
    assert worst < 1e-6


def _g(lcg):
    u1 = max(lcg.next(), 1e-300)
    u2 = lcg.next()
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def _pr(inl, truth):
    tp = int(np.sum(inl & truth))
    fp = int(np.sum(inl & ~truth))
    return tp / max(tp + fp, 1), tp / truth.sum(), tp + fp


# matching.rs:1237-1265
def test_ransac_separates_outliers():
    for out, seed in [(0.3, 5), (0.5, 9)]:
        x1, x2, truth, c1, c2 = ob.correspondences(300, 0.5, out, seed)
        F, inl, _ = ob.ransac_fundamental(x1, x2)
        prec, rec, _ = _pr(inl, truth)
        rms = np.sqrt(np.mean([ob.sampson_error(F, x1[i], x2[i]) for i in np.flatnonzero(truth)]))
        assert prec >= 0.97 and rec >= 0.98 and rms < 0.6


def _many_seed_case(d):
    x1, x2, truth, c1, c2 = ob.correspondences(300, 0.5, 0.5, d * 7919)
    g = ob.fundamental_from_cameras(c1, c2)
    gi = np.array([ob.sampson_error(g, x1[i], x2[i]) < TH2 for i in range(300)])
    gt = _pr(gi, truth)
    out = []
    for r in range(1, 6):
        res = ob.ransac_fundamental(x1, x2, ob.default_cfg(seed=r))
        out.append((d, r, gt, None if res is None else _pr(res[1], truth)))
    return out


@pytest.fixture(scope="module")
def many_seed_cases():
    # ctypes drops the GIL during each call, so threads run the oracle in parallel.
    with ThreadPoolExecutor(8) as ex:
        return [c for chunk in ex.map(_many_seed_case, range(1, 101)) for c in chunk]


# matching.rs:1306-1321, :1323-1347, :1349-1367, :1369-1387
def test_ransac_many_seeds(many_seed_cases):
    cases = many_seed_cases
    assert len(cases) == 500
    assert sum(c[3] is None for c in cases) == 0
    bad_prec = [c for c in cases if c[3][0] < c[2][0] - 0.04]
    assert not bad_prec
    bad_rec = [c for c in cases
               if not (c[3][1] >= min(c[2][1] - 0.01, 0.98)
                       or (c[3][2] >= c[2][2] and c[3][1] >= c[2][1] - 0.03))]
    assert not bad_rec
    bad_search = [c for c in cases if c[3][2] < 0.98 * c[2][2]]
    assert not bad_search


# matching.rs:2151-2186
def test_ransac_4000_correspondences_lm_stops_early():
    x1, x2, _, c1, c2 = ob.correspondences(4000, 0.5, 0.5, 4242)
    g = ob.fundamental_from_cameras(c1, c2)
    gt = sum(ob.sampson_error(g, x1[i], x2[i]) < TH2 for i in range(4000))
    F, inl, _ = ob.ransac_fundamental(x1, x2)
    st = ob.ransac_last_stats()
    assert inl.sum() >= 0.98 * gt
    assert st["lm_calls"] > 0
    assert 2 * st["lm_converged"] > st["lm_calls"]
    assert st["lm_stalled_at_start"] == 0
    # matching.rs:2189-2190: about 1710 hypotheses for this input.
    assert abs(st["iters"] - 1710) <= 20


# matching.rs:1747-1773
def test_degenerate_inputs_are_rejected_without_panic():
    x1, x2, _, _, _ = ob.correspondences(50, 0.0, 0.0, 3)
    assert ob.fundamental_8pt(x1[:7], x2[:7]) is None
    assert ob.ransac_fundamental(x1[:7], x2[:7]) is None
    assert ob.fundamental_8pt(x1, x2[:40]) is None
    assert ob.ransac_fundamental(x1, x2[:40]) is None
    for bad in (np.nan, np.inf):
        y = x2.copy()
        y[17, 0] = bad
        assert ob.fundamental_8pt(x1, y) is None
        assert ob.ransac_fundamental(x1, y) is None
    i = np.arange(60, dtype=float)
    l1 = np.stack([100 + 13 * i, 200 + 3 * i], 1)
    l2 = np.stack([100 + 13 * i, 350 - 2 * i], 1)
    assert ob.fundamental_8pt(l1, l2) is None
    assert ob.ransac_fundamental(l1, l2) is None


# matching.rs:1808-1822
def test_noisy_collinear_points_are_rejected():
    for sigma in (0.5, 1.0):
        for seed in range(1, 6):
            l1, l2 = ob.noisy_lines(sigma, seed)
            assert ob.fundamental_8pt(l1, l2) is None
            assert ob.ransac_fundamental(l1, l2) is None


# matching.rs:1902-1923
def test_random_correspondences_are_not_confirmed():
    for seed in range(1, 11):
        g = ob.Lcg(seed)
        pts = np.array([[g.next() * 960.0, g.next() * 540.0] for _ in range(400)])
        x1, x2 = pts[:200], pts[200:]
        assert ob.ransac_fundamental(x1, x2) is None
        loose = ob.ransac_fundamental(x1, x2, ob.default_cfg(min_inlier_ratio=0.0))
        cnt = 0 if loose is None else int(loose[1].sum())
        assert cnt < 40


# matching.rs:2274-2304
def test_non_overlapping_pairs_are_rejected():
    legacy = ob.default_cfg(min_inliers=8, max_chance_prob=0.0)
    rejected = legacy_pass = 0
    for seed in range(1, 201):
        g = ob.Lcg(seed * 7919)
        n = 15 + int(g.next() * 26.0)
        pts = np.array([[g.next() * 960.0, g.next() * 540.0] for _ in range(2 * n)])
        x1, x2 = pts[:n], pts[n:]
        rejected += ob.ransac_fundamental(x1, x2) is None
        legacy_pass += ob.ransac_fundamental(x1, x2, legacy) is not None
    assert rejected == 200
    assert legacy_pass > 0


# matching.rs:2306-2325
def test_small_true_overlap_is_still_confirmed():
    for seed in range(1, 21):
        x1, x2, truth, _, _ = ob.correspondences(30, 0.5, 1.0 / 3.0, seed)
        nt = int(truth.sum())
        if nt < 16:
            continue
        r = ob.ransac_fundamental(x1, x2)
        assert r is not None, seed
        assert (r[1] & truth).sum() >= 0.9 * nt


# matching.rs:2232-2263
def test_line_plus_few_general_points_is_not_confirmed():
    for k in (2, 4, 6, 10):
        confirmed = 0
        for seed in range(1, 11):
            x1, x2, h1, h2 = ob.line_plus_general(k, 0.5, seed)
            r = ob.ransac_fundamental(x1, x2)
            if r is None:
                continue
            confirmed += 1
            e = sorted(ob.sampson_error(r[0], p, q) ** 0.5 for p, q in zip(h1, h2))
            assert e[len(e) // 2] <= 3.0, (k, seed)
        if k == 10:
            assert confirmed == 10


# matching.rs:1847-1864
def test_homography_support_separates_planar_scene():
    for seed in range(1, 6):
        x1, x2 = ob.planar_correspondences(200, 0.5, seed)
        assert ob.homography_support(x1, x2, 1.5) >= 150
        x1, x2, _, _, _ = ob.correspondences(200, 0.5, 0.0, seed)
        assert ob.homography_support(x1, x2, 1.5) < 100


# matching.rs:2327-2343
def test_homography_support_finds_plane_among_outliers():
    for seed in range(1, 6):
        x1, x2 = ob.planar_correspondences(140, 0.5, seed)
        g = ob.Lcg(seed + 77)
        o = np.array([[g.next() * 960.0, g.next() * 540.0, g.next() * 960.0, g.next() * 540.0]
                      for _ in range(60)])
        x1 = np.concatenate([x1, o[:, :2]])
        x2 = np.concatenate([x2, o[:, 2:]])
        assert ob.homography_support(x1, x2, 1.5) >= 100


# matching.rs:1954-1976
def test_planar_scene_selects_homography():
    c1, c2 = ob.two_cameras()
    f0 = ob.fundamental_from_cameras(c1, c2)
    for seed in range(1, 6):
        x1, x2 = ob.planar_correspondences(200, 0.5, seed)
        F, inl, _ = ob.ransac_fundamental(x1, x2)
        m = ob.select_two_view_model(x1[inl], x2[inl], F, 0.5)
        assert m.model == 1
        assert ob.select_two_view_model(x1, x2, f0, 0.5).model == 1


# matching.rs:1978-2011
def test_model_selection_robust_to_underestimated_sigma():
    c1, c2 = ob.two_cameras()
    f0 = ob.fundamental_from_cameras(c1, c2)
    for actual in (0.5, 1.0, 1.5):
        for seed in range(1, 6):
            x1, x2 = ob.planar_correspondences(200, actual, seed)
            m = ob.select_two_view_model(x1, x2, f0, 0.5)
            assert m.model == 1
            assert 0.7 * actual < m.sigma_px < 1.4 * max(actual, 0.5)
    for seed in range(1, 6):
        x1, x2 = ob.ground_with_buildings(300, 0.05, 0.5, seed)
        assert ob.select_two_view_model(x1, x2, f0, 0.5).model == 0


# matching.rs:2013-2027
def test_model_selection_rejects_non_finite_sigma():
    c1, c2 = ob.two_cameras()
    f0 = ob.fundamental_from_cameras(c1, c2)
    x1, x2 = ob.planar_correspondences(50, 0.5, 1)
    for s in (np.inf, -np.inf, np.nan, 0.0, -1.0):
        assert ob.select_two_view_model(x1, x2, f0, s) is None
    l1 = np.array([[10.0 * i, 5.0 * i] for i in range(20)])
    l2 = l1 + np.array([3.0, 1.0])
    assert ob.select_two_view_model(l1, l2, f0, np.inf) is None


# matching.rs:2029-2072
def test_ground_scene_with_few_buildings_keeps_fundamental():
    c1, c2 = ob.two_cameras()
    f0 = ob.fundamental_from_cameras(c1, c2)
    for bld in (0.05, 0.1):
        for seed in range(1, 6):
            x1, x2 = ob.ground_with_buildings(300, bld, 0.5, seed)
            assert ob.select_two_view_model(x1, x2, f0, 0.5).model == 0
            F, inl, _ = ob.ransac_fundamental(x1, x2)
            assert ob.select_two_view_model(x1[inl], x2[inl], F, 0.5).model == 0
            truth = np.array([ob.sampson_error(f0, x1[i], x2[i]) < TH2 for i in range(300)])
            assert (truth & inl).sum() >= 0.95 * truth.sum()
            nb = int(round(300 * bld))
            ok = sum(ob.sampson_error(F, x1[i], x2[i]) < TH2 for i in range(nb))
            assert ok >= 0.9 * nb


# matching.rs:2074-2084
def test_general_scene_selects_fundamental():
    c1, c2 = ob.two_cameras()
    f0 = ob.fundamental_from_cameras(c1, c2)
    for seed in range(1, 6):
        x1, x2, _, _, _ = ob.correspondences(200, 0.5, 0.0, seed)
        m = ob.select_two_view_model(x1, x2, f0, 0.5)
        assert m.model == 0 and m.gric_f < m.gric_h


# matching.rs:2821-2849
def test_model_selection_rejects_non_finite_coordinates():
    c1, c2 = ob.two_cameras()
    f0 = ob.fundamental_from_cameras(c1, c2)
    x1, x2, _, _, _ = ob.correspondences(50, 0.5, 0.0, 1)
    assert ob.select_two_view_model(x1, x2, f0, 0.5).model == 0
    for bad in (np.nan, np.inf, -np.inf):
        for which, i in [(0, 3), (1, 3), (0, 49), (1, 0)]:
            y1, y2 = x1.copy(), x2.copy()
            if which == 0:
                y1[i, 0] = bad
            else:
                y2[i, 1] = bad
            assert ob.select_two_view_model(y1, y2, f0, 0.5) is None
        fb = f0.copy()
        fb[1, 2] = bad
        assert ob.select_two_view_model(x1, x2, fb, 0.5) is None
