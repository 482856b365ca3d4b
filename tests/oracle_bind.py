"""ctypes bindings of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP library against it. It is
never imported by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_sz = ctypes.c_size_t
_d = ctypes.c_double


class RansacCfg(ctypes.Structure):
    _fields_ = [("threshold_px", _d), ("min_inlier_ratio", _d), ("max_iters", _u64),
                ("confidence", _d), ("seed", _u64), ("min_inliers", _u64),
                ("max_chance_prob", _d)]


class ModelSel(ctypes.Structure):
    _fields_ = [("has_selection", ctypes.c_int), ("model", ctypes.c_int), ("gric_f", _d),
                ("gric_h", _d), ("parallax", _u64), ("parallax_needed", _u64),
                ("sigma_px", _d)]


_SIGS = {
    "orc_lcg_next": (_d, [_vp]),
    "orc_match_fixture": (ctypes.c_long, [_u64, _sz, _sz, _vp, _vp, _sz]),
    "orc_match_fixture_nb": (_sz, [_sz, _sz]),
    "orc_feature": (None, [_vp, _vp]),
    "orc_sq_dist": (ctypes.c_float, [_vp, _vp]),
    "orc_sq_dist_spec": (ctypes.c_float, [_vp, _vp]),
    "orc_sq_dist_seq": (ctypes.c_float, [_vp, _vp]),
    "orc_ratio_match": (_sz, [_vp, _sz, _vp, _sz, ctypes.c_float, ctypes.c_int, _vp]),
    "orc_ratio_match_brute": (_sz, [_vp, _sz, _vp, _sz, ctypes.c_float, ctypes.c_int, _vp]),
    "orc_ratio_match_state": (None, [_vp, _sz, _vp, _sz, _vp, _vp, _vp, _vp, _vp]),
    "orc_pairs_hash": (_u64, [_vp, _sz]),
    "orc_candidate_pairs": (_sz, [_vp, _sz, _u64, _u64, _u64, _vp, _sz]),
    "orc_scheduled_pairs": (_sz, [_vp, _sz, _vp, _vp, _sz]),
    "orc_verified_matches": (ctypes.c_long, [_vp, _sz, _vp, _sz, _vp]),
    "orc_detector_cfg_default": (None, [_vp]),
    "orc_detect_and_describe": (ctypes.c_long, [_vp, _sz, _sz, _vp, _vp, _vp, _sz]),
    "orc_gaussian_blur": (ctypes.c_int, [_vp, _sz, _sz, ctypes.c_float, _vp]),
    "orc_gaussian_taps": (ctypes.c_long, [ctypes.c_float, _vp, _sz]),
    "orc_upsample2": (None, [_vp, _sz, _sz, _vp]),
    "orc_downsample": (None, [_vp, _sz, _sz, _vp]),
    "orc_dominant_orientations": (_sz, [_vp, _sz, _sz, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, _vp]),
    "orc_describe": (None, [_vp, _sz, _sz, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                            ctypes.c_float, _vp]),
    "orc_expf": (ctypes.c_float, [ctypes.c_float]),
    "orc_atan2f": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
    "orc_powf": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
    "orc_sincosf": (None, [ctypes.c_float, _vp, _vp]),
}

_lib = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build_oracle()
        L = ctypes.CDLL(ORACLE_LIB)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def P(a: np.ndarray):
    return a.ctypes.data if a.size else None


class Lcg:
    """matching.rs:1141-1154 (through the oracle so every draw is the same)."""

    def __init__(self, seed: int):
        self.state = np.array([seed], dtype=np.uint64)

    def next(self) -> float:
        return lib().orc_lcg_next(P(self.state))

    def feature(self) -> np.ndarray:
        d = np.zeros(128, np.float32)
        lib().orc_feature(P(self.state), P(d))
        return d


def match_fixture(seed: int, na: int, extra: int):
    L = lib()
    nb = L.orc_match_fixture_nb(na, extra)
    a = np.zeros((na, 128), np.float32)
    b = np.zeros((nb, 128), np.float32)
    r = L.orc_match_fixture(seed, na, extra, P(a) if na else None, P(b) if nb else None, nb)
    assert r == nb
    return a, b


def _c(x):
    return np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 128)


def ratio_match(a, b, ratio, mutual):
    a, b = _c(a), _c(b)
    out = np.zeros((max(1, a.shape[0]), 2), np.uint64)
    n = lib().orc_ratio_match(P(a), a.shape[0], P(b), b.shape[0], ratio, int(mutual), P(out))
    return [(int(i), int(j)) for i, j in out[:n]]


def ratio_match_brute(a, b, ratio, mutual):
    a, b = _c(a), _c(b)
    out = np.zeros((max(1, a.shape[0]), 2), np.uint64)
    n = lib().orc_ratio_match_brute(P(a), a.shape[0], P(b), b.shape[0], ratio, int(mutual), P(out))
    return [(int(i), int(j)) for i, j in out[:n]]


def ratio_match_state(a, b):
    a, b = _c(a), _c(b)
    na, nb = a.shape[0], b.shape[0]
    j1 = np.zeros(na, np.uint64)
    d1 = np.zeros(na, np.float32)
    d2 = np.zeros(na, np.float32)
    cd = np.zeros(nb, np.float32)
    ci = np.zeros(nb, np.uint64)
    lib().orc_ratio_match_state(P(a), na, P(b), nb, P(j1), P(d1), P(d2), P(cd), P(ci))
    return j1, d1, d2, ci, cd


def sq_dist(p, r, which="impl"):
    p = np.ascontiguousarray(p, np.float32)
    r = np.ascontiguousarray(r, np.float32)
    f = {"impl": lib().orc_sq_dist, "spec": lib().orc_sq_dist_spec,
         "seq": lib().orc_sq_dist_seq}[which]
    return np.float32(f(P(p), P(r)))


def pairs_hash(m) -> int:
    arr = np.ascontiguousarray(np.asarray(m, dtype=np.uint64).reshape(-1, 2))
    return int(lib().orc_pairs_hash(P(arr), arr.shape[0]))


def candidate_pairs(views, temporal, cross, pow2_max):
    v = np.ascontiguousarray(np.asarray(views, dtype=np.uint64).reshape(-1, 2))
    mx = 0xFFFFFFFFFFFFFFFF
    n = lib().orc_candidate_pairs(P(v), v.shape[0], min(temporal, mx), min(cross, mx),
                                  min(pow2_max, mx), None, 0)
    out = np.zeros((max(1, n), 2), np.uint64)
    lib().orc_candidate_pairs(P(v), v.shape[0], min(temporal, mx), min(cross, mx),
                              min(pow2_max, mx), P(out), n)
    return [(int(i), int(j)) for i, j in out[:n]]


# ---------------------------------------------------------------- RANSAC
_SIGS_R = {
    "orc_ransac_cfg_default": (None, [_vp]),
    "orc_adaptive_iterations": (_u64, [_d, ctypes.c_int32, _d, _u64]),
    "orc_binomial_tail": (_d, [_u64, _u64, _d]),
    "orc_chance_inlier_probability": (_d, [_u64, _u64, _d]),
    "orc_epipolar_band_probability": (_d, [_vp, _sz, _d]),
    "orc_parallax_needed": (_u64, [_u64]),
    "orc_fundamental_8pt": (ctypes.c_int, [_vp, _vp, _sz, _vp]),
    "orc_sampson_error": (_d, [_vp, _vp, _vp]),
    "orc_sampson_residual_jacobian": (None, [_vp, _d, _d, _vp, _vp, _vp, _vp]),
    "orc_normalizer": (None, [_vp, _sz, _vp]),
    "orc_refine_sampson": (ctypes.c_int, [_vp, _vp, _vp, _sz, _sz, _vp, _vp, _vp]),
    "orc_homography_support": (_u64, [_vp, _vp, _sz, _d]),
    "orc_homography_dlt": (ctypes.c_int, [_vp, _vp, _sz, _vp]),
    "orc_homography_4pt": (ctypes.c_int, [_vp, _vp, _vp]),
    "orc_select_two_view_model": (ctypes.c_int, [_vp, _vp, _sz, _vp, _d, _vp]),
    "orc_ransac_fundamental": (ctypes.c_int, [_vp, _vp, _sz, _vp, _vp, _vp, _vp]),
    "orc_ransac_last_stats": (None, [_vp, _vp, _vp, _vp, _vp]),
    "orc_ransac_trace_begin": (None, [_vp, _sz, _vp, _sz]),
    "orc_ransac_trace_end": (_sz, []),
    "orc_eigen_min9": (None, [_vp, _vp, _vp]),
    "orc_eigen_min_inv9": (ctypes.c_int, [_vp, _vp]),
    "orc_two_cameras": (None, [_vp, _vp]),
    "orc_project": (ctypes.c_int, [_vp, _vp, _vp]),
    "orc_fundamental_from_cameras": (None, [_vp, _vp, _vp]),
    "orc_correspondences": (None, [_sz, _d, _d, _u64, _vp, _vp, _vp]),
    "orc_planar_correspondences": (None, [_sz, _d, _u64, _vp, _vp]),
    "orc_ground_with_buildings": (None, [_sz, _d, _d, _u64, _vp, _vp]),
    "orc_noisy_lines": (None, [_d, _u64, _vp, _vp]),
    "orc_line_plus_general": (None, [_sz, _d, _u64, _vp, _vp, _vp, _vp]),
}
_SIGS.update(_SIGS_R)


def default_cfg(**kw) -> RansacCfg:
    c = RansacCfg()
    lib().orc_ransac_cfg_default(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _xy(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, 2))


def adaptive_iterations(w, s, conf, mx):
    return int(lib().orc_adaptive_iterations(w, s, conf, mx))


def binomial_tail(n, k, p):
    return float(lib().orc_binomial_tail(n, k, p))


def parallax_needed(n):
    return int(lib().orc_parallax_needed(n))


def fundamental_8pt(x1, x2):
    x1, x2 = _xy(x1), _xy(x2)
    F = np.zeros(9)
    n = min(x1.shape[0], x2.shape[0]) if x1.shape[0] == x2.shape[0] else 0
    ok = lib().orc_fundamental_8pt(P(x1), P(x2), n, P(F)) if x1.shape[0] == x2.shape[0] else 0
    return F.reshape(3, 3) if ok else None


def sampson_error(F, p, q):
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    p = np.ascontiguousarray(p, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    return float(lib().orc_sampson_error(P(F), P(p), P(q)))


def sampson_residual_jacobian(g, s1, s2, a, b):
    g = np.ascontiguousarray(g, np.float64).reshape(9)
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    r = np.zeros(1)
    j = np.zeros(9)
    lib().orc_sampson_residual_jacobian(P(g), s1, s2, P(a), P(b), P(r), P(j))
    return float(r[0]), j


def normalizer(p):
    p = _xy(p)
    T = np.zeros(9)
    lib().orc_normalizer(P(p), p.shape[0], P(T))
    return T.reshape(3, 3)


def homography_dlt(x1, x2):
    x1, x2 = _xy(x1), _xy(x2)
    H = np.zeros(9)
    ok = lib().orc_homography_dlt(P(x1), P(x2), x1.shape[0], P(H))
    return H.reshape(3, 3) if ok else None


def homography_4pt(x1, x2):
    x1, x2 = _xy(x1), _xy(x2)
    assert x1.shape[0] == 4 and x2.shape[0] == 4
    H = np.zeros(9)
    ok = lib().orc_homography_4pt(P(x1), P(x2), P(H))
    return H.reshape(3, 3) if ok else None


def homography_support(x1, x2, th):
    x1, x2 = _xy(x1), _xy(x2)
    return int(lib().orc_homography_support(P(x1), P(x2), x1.shape[0], th))


def select_two_view_model(x1, x2, F, sigma):
    x1, x2 = _xy(x1), _xy(x2)
    if x1.shape[0] != x2.shape[0]:
        return None
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    m = ModelSel()
    ok = lib().orc_select_two_view_model(P(x1), P(x2), x1.shape[0], P(F), sigma, ctypes.byref(m))
    return m if ok else None


def ransac_fundamental(x1, x2, cfg=None):
    """(F 3x3, inliers bool[n], ModelSel) or None."""
    x1, x2 = _xy(x1), _xy(x2)
    if x1.shape[0] != x2.shape[0]:
        return None
    cfg = cfg or default_cfg()
    n = x1.shape[0]
    F = np.zeros(9)
    inl = np.zeros(max(n, 1), np.uint8)
    sel = ModelSel()
    ok = lib().orc_ransac_fundamental(P(x1), P(x2), n, ctypes.byref(cfg), P(F), P(inl),
                                      ctypes.byref(sel))
    if not ok:
        return None
    return F.reshape(3, 3), inl[:n].astype(bool), sel


def ransac_last_stats():
    v = [np.zeros(1, np.uint64) for _ in range(5)]
    lib().orc_ransac_last_stats(*[P(x) for x in v])
    return dict(zip(["iters", "lo_runs", "lm_calls", "lm_converged", "lm_stalled_at_start"],
                    [int(x[0]) for x in v]))


def eigen_min_inv9(A):
    """(degenerate, unit eigenvector of the smallest eigenvalue or None)."""
    A = np.ascontiguousarray(A, np.float64).reshape(81)
    vec = np.zeros(9)
    deg = lib().orc_eigen_min_inv9(P(A), P(vec))
    return bool(deg), (None if deg else vec)


def eigen_min9(A):
    A = np.ascontiguousarray(A, np.float64).reshape(81)
    ev = np.zeros(9)
    vec = np.zeros(9)
    lib().orc_eigen_min9(P(A), P(ev), P(vec))
    return ev, vec


def two_cameras():
    c1 = np.zeros(18)
    c2 = np.zeros(18)
    lib().orc_two_cameras(P(c1), P(c2))
    return c1, c2


def fundamental_from_cameras(c1, c2):
    F = np.zeros(9)
    lib().orc_fundamental_from_cameras(P(np.ascontiguousarray(c1)), P(np.ascontiguousarray(c2)),
                                       P(F))
    return F.reshape(3, 3)


def correspondences(n, sigma, out, seed):
    x1 = np.zeros((n, 2))
    x2 = np.zeros((n, 2))
    t = np.zeros(n, np.uint8)
    lib().orc_correspondences(n, sigma, out, seed, P(x1), P(x2), P(t))
    c1, c2 = two_cameras()
    return x1, x2, t.astype(bool), c1, c2


def planar_correspondences(n, sigma, seed):
    x1 = np.zeros((n, 2))
    x2 = np.zeros((n, 2))
    lib().orc_planar_correspondences(n, sigma, seed, P(x1), P(x2))
    return x1, x2


def ground_with_buildings(n, bld, sigma, seed):
    x1 = np.zeros((n, 2))
    x2 = np.zeros((n, 2))
    lib().orc_ground_with_buildings(n, bld, sigma, seed, P(x1), P(x2))
    return x1, x2


def noisy_lines(sigma, seed):
    l1 = np.zeros((60, 2))
    l2 = np.zeros((60, 2))
    lib().orc_noisy_lines(sigma, seed, P(l1), P(l2))
    return l1, l2


def line_plus_general(k, sigma, seed):
    x1 = np.zeros((60 + k, 2))
    x2 = np.zeros((60 + k, 2))
    h1 = np.zeros((100, 2))
    h2 = np.zeros((100, 2))
    lib().orc_line_plus_general(k, sigma, seed, P(x1), P(x2), P(h1), P(h2))
    return x1, x2, h1, h2


class TraceRec(ctypes.Structure):
    """orc_trace_rec: one iteration of the hypothesis loop (matching.rs:967-1036)."""
    _fields_ = [("it", _u64), ("idx", _u64 * 8), ("ok", ctypes.c_int32), ("lo", ctypes.c_int32),
                ("raw_cnt", _u64), ("lo_cnt", _u64), ("best_before", _u64),
                ("iters_after", _u64), ("F", _d * 9), ("F_lo", _d * 9),
                ("npool", ctypes.c_int32), ("pad", ctypes.c_int32), ("pool_cnt", _u64 * 5),
                ("pool_origin", _u64 * 5), ("stage_ns", _u64 * 6), ("stage_ok", ctypes.c_int32 * 6),
                ("stage_cnt", _u64 * 6), ("stage_F", (_d * 9) * 6)]


def ransac_traced(x1, x2, cfg=None, cap=4096):
    """(result as ransac_fundamental, trace records, LO masks (records x n bool))."""
    x1, x2 = _xy(x1), _xy(x2)
    n = x1.shape[0]
    recs = (TraceRec * cap)()
    masks = np.zeros((cap, max(n, 1)), np.uint8)
    lib().orc_ransac_trace_begin(ctypes.addressof(recs), cap, P(masks), n)
    try:
        res = ransac_fundamental(x1, x2, cfg)
    finally:
        k = lib().orc_ransac_trace_end()
    return res, recs[:k], masks[:k, :n].astype(bool)


class PairSched(ctypes.Structure):
    _fields_ = [("temporal", _u64), ("pow2_max", _u64), ("cross_kind", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("range", _u64), ("right_min", _u64), ("left_min", _u64),
                ("max", _u64), ("step", _u64)]


def pair_schedule(kind="formation", **kw):
    """PairSchedule::default() (formation) or ::spec(), fields overridable (matching.rs:73-99)."""
    s = PairSched(temporal=5, pow2_max=16, cross_kind=1, range=4, right_min=20, left_min=20,
                  max=40, step=4)
    if kind == "spec":
        s.cross_kind = 0
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def scheduled_pairs(views, sch):
    v = np.ascontiguousarray(np.asarray(views, dtype=np.uint64).reshape(-1, 2))
    n = lib().orc_scheduled_pairs(P(v), v.shape[0], ctypes.byref(sch), None, 0)
    out = np.zeros((max(1, n), 2), np.uint64)
    lib().orc_scheduled_pairs(P(v), v.shape[0], ctypes.byref(sch), P(out), n)
    return [(int(i), int(j)) for i, j in out[:n]]


def verified_matches(matches, inliers):
    """list, or ("err", (len(matches), len(inliers)))."""
    m = np.ascontiguousarray(np.asarray(matches, dtype=np.uint64).reshape(-1, 2))
    f = np.ascontiguousarray(np.asarray(inliers, dtype=np.uint8).reshape(-1))
    out = np.zeros((max(1, m.shape[0]), 2), np.uint64)
    k = lib().orc_verified_matches(P(m), m.shape[0], P(f), f.shape[0], P(out))
    if k < 0:
        return ("err", (m.shape[0], f.shape[0]))
    return [(int(i), int(j)) for i, j in out[:k]]


# ---- feature detector (features.rs, oracle/features_oracle.c) ----

class DetCfg(ctypes.Structure):
    """DetectorConfig (features.rs:259-289)."""
    _fields_ = [("octaves", _u64), ("scales", _u64), ("sigma0", ctypes.c_float),
                ("contrast", ctypes.c_float), ("edge_ratio", ctypes.c_float),
                ("upscale_fill", ctypes.c_uint32), ("max_features", _u64),
                ("upscale_below", _u64)]


def detector_cfg(**kw) -> DetCfg:
    c = DetCfg()
    lib().orc_detector_cfg_default(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _f32img(img):
    return np.ascontiguousarray(img, dtype=np.float32)


def detect_and_describe(img, cfg=None, describe=True):
    """(keypoints (n, 5) float32 [x, y, sigma, response, angle], descriptors (n, 128))."""
    img = _f32img(img)
    cfg = cfg or detector_cfg()
    h, w = img.shape
    n = lib().orc_detect_and_describe(P(img), w, h, ctypes.byref(cfg), None, None, 0)
    assert n >= 0, "oracle allocation failure"
    kps = np.zeros((max(n, 1), 5), np.float32)
    desc = np.zeros((max(n, 1), 128), np.float32)
    lib().orc_detect_and_describe(P(img), w, h, ctypes.byref(cfg), P(kps), P(desc), n)
    return kps[:n], (desc[:n] if describe else None)


def gaussian_blur(img, sigma):
    img = _f32img(img)
    out = np.zeros_like(img)
    assert lib().orc_gaussian_blur(P(img), img.shape[1], img.shape[0], ctypes.c_float(sigma),
                                   P(out)) == 0
    return out


def gaussian_taps(sigma):
    k = np.zeros(1025, np.float32)
    r = lib().orc_gaussian_taps(ctypes.c_float(sigma), P(k), 1025)
    return k[:2 * r + 1]


def upsample2(img):
    img = _f32img(img)
    out = np.zeros((2 * img.shape[0], 2 * img.shape[1]), np.float32)
    lib().orc_upsample2(P(img), img.shape[1], img.shape[0], P(out))
    return out


def dominant_orientations(img, x, y, sigma):
    img = _f32img(img)
    out = np.zeros(36, np.float32)
    n = lib().orc_dominant_orientations(P(img), img.shape[1], img.shape[0], ctypes.c_float(x),
                                        ctypes.c_float(y), ctypes.c_float(sigma), P(out))
    return out[:n]


def describe(img, x, y, sigma, angle):
    img = _f32img(img)
    out = np.zeros(128, np.float32)
    lib().orc_describe(P(img), img.shape[1], img.shape[0], ctypes.c_float(x), ctypes.c_float(y),
                       ctypes.c_float(sigma), ctypes.c_float(angle), P(out))
    return out


def mathf(name, *args):
    """orc_expf / orc_atan2f / orc_powf, or orc_sincosf -> (sin, cos)."""
    if name == "sincos":
        s, c = ctypes.c_float(), ctypes.c_float()
        lib().orc_sincosf(ctypes.c_float(args[0]), ctypes.byref(s), ctypes.byref(c))
        return s.value, c.value
    return getattr(lib(), "orc_" + name)(*[ctypes.c_float(a) for a in args])


# image generators of the reference's detector tests (features.rs:823-985),
# in float32 numpy (their last bits need not equal the reference's: the tests
# using them are statistical, and GPU parity feeds both sides the same image)

def blob(w, h, cx, cy, s):
    """features.rs:823-832."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    r2 = (x - np.float32(cx)) ** 2 + (y - np.float32(cy)) ** 2
    return (np.float32(0.2) + np.float32(0.6) * np.exp(-r2 / np.float32(2.0 * s * s))).astype(np.float32)


def _lcg_stream(seed):
    st = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        st = (st * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        yield np.float32(st >> 33) / np.float32(1 << 31)


def texture(w, h, seed):
    """features.rs:915-943: 400 random Gaussian blobs on 0.5."""
    g = _lcg_stream(seed)
    blobs = [[next(g) * np.float32(400.0), next(g) * np.float32(400.0),
              np.float32(1.5) + next(g) * np.float32(5.0), next(g) - np.float32(0.5)]
             for _ in range(400)]
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    px = x - np.float32(w / 2.0) + np.float32(200.0)
    py = y - np.float32(h / 2.0) + np.float32(200.0)
    v = np.full((h, w), 0.5, np.float32)
    for bx, by, bs, ba in blobs:
        d2 = (px - bx) ** 2 + (py - by) ** 2
        m = d2 < np.float32(16.0) * bs * bs
        v[m] += np.float32(0.5) * ba * np.exp(-d2[m] / (np.float32(2.0) * bs * bs))
    return v


_TEX_CACHE = {}


def texture_warped(w, h, seed, th, sc):
    """features.rs:946-966: the 800 x 800 texture rotated by th, scaled by sc."""
    if seed not in _TEX_CACHE:
        _TEX_CACHE[seed] = texture(800, 800, seed)
    big = _TEX_CACHE[seed]
    c, s = np.float32(np.cos(th)), np.float32(np.sin(th))
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    fx, fy = x - np.float32(w / 2.0), y - np.float32(h / 2.0)
    ox = (c * fx + s * fy) / np.float32(sc) + np.float32(400.0)
    oy = (-s * fx + c * fy) / np.float32(sc) + np.float32(400.0)
    x0 = np.floor(ox).astype(np.int64)
    y0 = np.floor(oy).astype(np.int64)
    ax, ay = ox - x0, oy - y0

    def at(xx, yy):
        return big[np.minimum(yy, 799), np.minimum(xx, 799)]
    return ((1 - ay) * ((1 - ax) * at(x0, y0) + ax * at(x0 + 1, y0))
            + ay * ((1 - ax) * at(x0, y0 + 1) + ax * at(x0 + 1, y0 + 1))).astype(np.float32)


def scene_texture(w, h, seed, n_blobs=None):
    """workloads.textured_image: the detector's textured test image."""
    import workloads
    return workloads.textured_image(w, h, seed, n_blobs)


# ---- multi-view tracks (tracks.rs, oracle/tracks_oracle.c) ----

class TrackCfg(ctypes.Structure):
    """TrackConfig (tracks.rs:109-125): policy 0 = Drop, 1 = Split."""
    _fields_ = [("policy", ctypes.c_int32), ("pad", ctypes.c_int32), ("min_length", _u64),
                ("max_tracks", _u64)]


TRACK_STATS = ("edges", "invalid_matches", "conflicts", "vote_cuts", "too_short", "inconsistent",
               "truncated", "tracks")


class TrackStats(ctypes.Structure):
    _fields_ = [(k, _u64) for k in TRACK_STATS]


_SIGS.update({
    "orc_build_tracks": (ctypes.c_long, [_sz, _vp, _vp, _vp, _vp, _sz, _vp, _vp, _vp, _vp, _sz,
                                         _vp, _sz, _vp]),
    "orc_displacement_outliers": (None, [_vp, _sz, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                         _vp]),
    "orc_select_tracks": (ctypes.c_long, [_vp, _vp, _sz, _sz, _vp]),
    "orc_set_displacement_order": (None, [ctypes.c_int]),
    "orc_displacement_offsets": (None, [_vp, _sz, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                        _vp, _vp]),
})


def track_cfg(policy="split", min_length=3, max_tracks=100_000):
    return TrackCfg(1 if policy == "split" else 0, 0, min_length, max_tracks)


def pack_track_input(pairs, keypoints):
    """Flat arrays of the C layout: pairs = [(image_a, image_b, [(fa, fb), ...]), ...],
    keypoints = [array (n_i, 2), ...]."""
    ia = np.array([p[0] for p in pairs], np.uint64).reshape(-1)
    ib = np.array([p[1] for p in pairs], np.uint64).reshape(-1)
    lens = [len(p[2]) for p in pairs]
    m_off = np.zeros(len(pairs) + 1, np.uint64)
    m_off[1:] = np.cumsum(lens, dtype=np.uint64) if pairs else []
    m = (np.concatenate([np.asarray(p[2], np.uint64).reshape(-1, 2) for p in pairs])
         if sum(lens) else np.zeros((0, 2), np.uint64))
    kl = [len(k) for k in keypoints]
    kp_off = np.zeros(len(keypoints) + 1, np.uint64)
    kp_off[1:] = np.cumsum(kl, dtype=np.uint64)
    kp = (np.concatenate([np.asarray(k, np.float64).reshape(-1, 2) for k in keypoints])
          if sum(kl) else np.zeros((0, 2), np.float64))
    return (np.ascontiguousarray(ia), np.ascontiguousarray(ib), m_off,
            np.ascontiguousarray(m, dtype=np.uint64), kp_off, np.ascontiguousarray(kp))


def build_tracks(pairs, keypoints, cfg=None):
    """(tracks as lists of (image, feature), stats dict)."""
    cfg = cfg or track_cfg()
    ia, ib, m_off, m, kp_off, kp = pack_track_input(pairs, keypoints)
    n = int(kp_off[-1])
    toff = np.zeros(n + 2, np.uint64)
    obs = np.zeros((max(1, n), 2), np.uint64)
    st = TrackStats()
    k = lib().orc_build_tracks(len(pairs), P(ia), P(ib), P(m_off), P(m), len(keypoints), P(kp_off),
                               P(kp), ctypes.byref(cfg), P(toff), toff.shape[0], P(obs), n,
                               ctypes.byref(st))
    assert k >= 0, "orc_build_tracks failed"
    tracks = [[(int(a), int(b)) for a, b in obs[toff[t]:toff[t + 1]]] for t in range(k)]
    return tracks, {f: int(getattr(st, f)) for f in TRACK_STATS}


def displacement_outliers(matches, kp_a, kp_b, floor=0.25, cap=3.0):
    m = np.ascontiguousarray(np.asarray(matches, np.uint64).reshape(-1, 2))
    a = np.ascontiguousarray(np.asarray(kp_a, np.float64).reshape(-1, 2))
    b = np.ascontiguousarray(np.asarray(kp_b, np.float64).reshape(-1, 2))
    out = np.zeros(max(1, m.shape[0]), np.uint8)
    lib().orc_displacement_outliers(P(m), m.shape[0], P(a), P(b), floor, cap, P(out))
    return out[:m.shape[0]].astype(bool)


def displacement_offsets(matches, kp_a, kp_b, floor=0.25, cap=3.0):
    """(flags, r, threshold) per match; r and threshold NAN where no fit ran."""
    m = np.ascontiguousarray(np.asarray(matches, np.uint64).reshape(-1, 2))
    a = np.ascontiguousarray(np.asarray(kp_a, np.float64).reshape(-1, 2))
    b = np.ascontiguousarray(np.asarray(kp_b, np.float64).reshape(-1, 2))
    out = np.zeros(max(1, m.shape[0]), np.uint8)
    offs = np.zeros((max(1, m.shape[0]), 2))
    lib().orc_displacement_offsets(P(m), m.shape[0], P(a), P(b), floor, cap, P(out), P(offs))
    n = m.shape[0]
    return out[:n].astype(bool), offs[:n, 0], offs[:n, 1]


class displacement_order:
    """Context manager: the oracle's displacement filter takes its 24 kept
    candidates in order `mode` (0 by (squared distance, index), 1 reversed,
    2 by index, 3 a seeded shuffle) inside the block."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        lib().orc_set_displacement_order(self.mode)
        return self

    def __exit__(self, *exc):
        lib().orc_set_displacement_order(0)
        return False


def select_tracks(tracks, max_tracks):
    """Indices of the selected tracks (lists of (image, feature)) in output order."""
    toff = np.zeros(len(tracks) + 1, np.uint64)
    toff[1:] = np.cumsum([len(t) for t in tracks])
    obs = np.ascontiguousarray(np.array([o for t in tracks for o in t], np.uint64).reshape(-1, 2))
    out = np.zeros(max(1, len(tracks)), np.uint64)
    k = lib().orc_select_tracks(P(toff), P(obs), len(tracks), max_tracks, P(out))
    return [int(x) for x in out[:k]]
