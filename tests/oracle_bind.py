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
