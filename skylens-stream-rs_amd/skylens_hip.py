"""Host-side mirror of skylens-stream's matching API over the libskylens_hip C-ABI.

The functions keep the reference's names, argument meaning and error
behaviour (crates/core/src/matching.rs):

* ``candidate_pairs(views, temporal, cross, pow2_max)``      -- :18-40
* ``ratio_match(a, b, ratio, mutual)``                       -- :167-215
* ``ransac_fundamental(x1, x2, cfg)`` -> ``FundamentalFit | None`` -- :941-1104

Descriptors are ``(n, 128) float32`` arrays (the gathered ``Feature.desc``
rows); coordinates are ``(n, 2) float64`` arrays. Degenerate inputs give the
reference's empty list / ``None``; invalid arguments raise ``ValueError``.

All compute runs in the HIP library. There is no CPU fallback: if the shared
library is missing or no gfx950 device is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os
import sys
import threading
import weakref
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libskylens_hip.so")

SKH_OK = 0
SKH_EINVAL = -1
SKH_EHIP = -2
SKH_ENOMEM = -3
SKH_ENODEV = -4

PRECISIONS = {"fp32": 0, "bf16": 1, "fp8": 2}
TILES = {"auto": -1, "128x64": 0, "128x128": 1, "256x64": 2, "256x128": 3}

# matching.rs:7-12
PAIR_TEMPORAL = 5
PAIR_CROSS = 4
PAIR_POW2_MAX = 16

_c_size_p = ctypes.POINTER(ctypes.c_size_t)
_vp = ctypes.c_void_p

# (name, restype, argtypes) of every symbol include/skylens_hip.h declares.
EXPORTS = [
    ("skh_precision_supported", ctypes.c_int, [ctypes.c_int]),
    ("skh_ctx_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    ("skh_ctx_destroy", ctypes.c_int, [_vp]),
    ("skh_ctx_set_tile", ctypes.c_int, [_vp, ctypes.c_int]),
    ("skh_ctx_stream", _vp, [_vp]),
    ("skh_ctx_last_error", ctypes.c_char_p, [_vp]),
    ("skh_candidate_pairs", ctypes.c_int,
     [_vp, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _vp,
      ctypes.c_size_t, _c_size_p]),
    ("skh_ratio_match", ctypes.c_int,
     [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, ctypes.c_float, ctypes.c_int,
      ctypes.c_int, _vp, ctypes.c_size_t, _c_size_p]),
    ("skh_ratio_match_dev", ctypes.c_int,
     [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, ctypes.c_float, ctypes.c_int,
      ctypes.c_int, _vp, _vp]),
    ("skh_ratio_match_dev_noalloc", ctypes.c_int,
     [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, ctypes.c_float, ctypes.c_int,
      ctypes.c_int, _vp, _vp]),
    ("skh_ratio_match_state", ctypes.c_int,
     [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, _vp, _vp, _vp, _vp, _vp]),
    ("skh_ransac_cfg_default", None, [_vp]),
    ("skh_ransac_fundamental", ctypes.c_int,
     [_vp, _vp, _vp, ctypes.c_size_t, _vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_int)]),
    ("skh_ctx_last_ransac_iters", ctypes.c_uint64, [_vp]),
    ("skh_plan_create", ctypes.c_int, [_vp, _vp, ctypes.POINTER(_vp)]),
    ("skh_plan_destroy", ctypes.c_int, [_vp]),
    ("skh_plan_run", ctypes.c_int, [_vp]),
    ("skh_plan_sync", ctypes.c_int, [_vp]),
    ("skh_plan_match_base", ctypes.c_int, [_vp, _vp]),
    ("skh_plan_results", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("skh_plan_stage_ms", ctypes.c_double, [_vp, ctypes.c_int]),
    ("skh_plan_set_timing", ctypes.c_int, [_vp, ctypes.c_int]),
    ("skh_plan_time_distance", ctypes.c_double, [_vp, ctypes.c_int]),
    ("skh_plan_ransac_profile", ctypes.c_int, [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_double)]),
    ("skh_plan_lowp_state", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("skh_ctx_autotune", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    ("skh_ctx_set_tile_table", ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    ("skh_ctx_tile_table", ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_int)]),
    ("skh_plan_tiles", ctypes.c_int, [_vp, _vp]),
    ("skh_pair_schedule_default", None, [_vp]),
    ("skh_pair_schedule_spec", None, [_vp]),
    ("skh_scheduled_pairs", ctypes.c_int, [_vp, ctypes.c_size_t, _vp, _vp, ctypes.c_size_t, _c_size_p]),
    ("skh_verified_matches", ctypes.c_int, [_vp, ctypes.c_size_t, _vp, ctypes.c_size_t, _vp, _c_size_p]),
    ("skh_plan_verified", ctypes.c_int, [_vp, _vp, _vp]),
    ("skh_detector_cfg_default", None, [_vp]),
    ("skh_detect_and_describe", ctypes.c_int,
     [_vp, _vp, ctypes.c_size_t, ctypes.c_size_t, _vp, _vp, _vp, ctypes.c_size_t, _c_size_p]),
    ("skh_detect_and_describe_dev", ctypes.c_int,
     [_vp, _vp, ctypes.c_size_t, ctypes.c_size_t, _vp, _vp, _vp, ctypes.c_size_t, _c_size_p]),
    ("skh_track_cfg_default", None, [_vp]),
    ("skh_build_tracks", ctypes.c_int,
     [_vp, ctypes.c_size_t, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp, _vp, _vp, _vp,
      ctypes.c_size_t, _vp, ctypes.c_size_t, _c_size_p, _vp]),
]

TILE_CLASSES = 25
TILE_NAMES = ["128x64", "128x128", "256x64", "256x128"]

RANSAC_PHASES = ["sampling", "hypotheses", "local_opt", "replay", "pool_lm", "final_gates",
                 "line_test", "model_selection", "batches", "lo_lanes", "total"]


class RansacConfig(ctypes.Structure):
    """matching.rs:762-793 (skh_ransac_cfg)."""
    _fields_ = [("threshold_px", ctypes.c_double), ("min_inlier_ratio", ctypes.c_double),
                ("max_iters", ctypes.c_uint64), ("confidence", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("min_inliers", ctypes.c_uint64),
                ("max_chance_prob", ctypes.c_double)]

    @classmethod
    def default(cls, **kw) -> "RansacConfig":
        c = cls()
        load_library().skh_ransac_cfg_default(ctypes.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class DetectorConfig(ctypes.Structure):
    """DetectorConfig (features.rs:259-289), skh_detector_cfg."""
    _fields_ = [("octaves", ctypes.c_uint64), ("scales", ctypes.c_uint64),
                ("sigma0", ctypes.c_float), ("contrast", ctypes.c_float),
                ("edge_ratio", ctypes.c_float), ("upscale_fill", ctypes.c_uint32),
                ("max_features", ctypes.c_uint64), ("upscale_below", ctypes.c_uint64)]

    @classmethod
    def default(cls, **kw) -> "DetectorConfig":
        c = cls()
        load_library().skh_detector_cfg_default(ctypes.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c


KEYPOINT_FIELDS = ("x", "y", "sigma", "response", "angle")  # features.rs:295-304

TRACKS_DROP, TRACKS_SPLIT = 0, 1  # ConflictPolicy (tracks.rs:103-107)


class TrackConfig(ctypes.Structure):
    """TrackConfig (tracks.rs:109-125), skh_track_cfg."""
    _fields_ = [("policy", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("min_length", ctypes.c_uint64), ("max_tracks", ctypes.c_uint64)]

    @classmethod
    def default(cls, **kw) -> "TrackConfig":
        c = cls()
        load_library().skh_track_cfg_default(ctypes.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c


TRACK_STATS = ("edges", "invalid_matches", "conflicts", "vote_cuts", "too_short", "inconsistent",
               "truncated", "tracks")


class TrackStats(ctypes.Structure):
    """TrackStats (tracks.rs:145-155), skh_track_stats."""
    _fields_ = [(k, ctypes.c_uint64) for k in TRACK_STATS]


class TrackInput:
    """build_tracks' inputs in the C-ABI's flat layout, with output buffers:
    pairs = [(image_a, image_b, [(fa, fb), ...]), ...] (PairMatches),
    keypoints = [(n_i, 2) pixels per image]."""

    def __init__(self, pairs, keypoints):
        self.n_pairs, self.n_images = len(pairs), len(keypoints)
        self.ia = np.ascontiguousarray([p[0] for p in pairs], dtype=np.uint64)
        self.ib = np.ascontiguousarray([p[1] for p in pairs], dtype=np.uint64)
        lens = [len(p[2]) for p in pairs]
        self.m_off = np.zeros(len(pairs) + 1, np.uint64)
        self.m_off[1:] = np.cumsum(lens)
        m = (np.concatenate([np.asarray(p[2], np.uint64).reshape(-1, 2) for p in pairs])
             if sum(lens) else np.zeros((0, 2), np.uint64))
        self.m = np.ascontiguousarray(m, dtype=np.uint64)
        kl = [len(k) for k in keypoints]
        self.kp_off = np.zeros(len(keypoints) + 1, np.uint64)
        self.kp_off[1:] = np.cumsum(kl)
        kp = (np.concatenate([np.asarray(k, np.float64).reshape(-1, 2) for k in keypoints])
              if sum(kl) else np.zeros((0, 2), np.float64))
        self.kp = np.ascontiguousarray(kp, dtype=np.float64)
        self.n = int(self.kp_off[-1])
        self.track_off = np.zeros(self.n + 2, np.uint64)
        self.obs = np.zeros((max(1, self.n), 2), np.uint64)


class PairSchedule(ctypes.Structure):
    """matching.rs:47-99 (skh_pair_schedule): cross_kind 0 = Spec { range },
    1 = Formation { right_min, left_min, max, step }."""
    _fields_ = [("temporal", ctypes.c_uint64), ("pow2_max", ctypes.c_uint64),
                ("cross_kind", ctypes.c_int32), ("pad", ctypes.c_int32), ("range", ctypes.c_uint64),
                ("right_min", ctypes.c_uint64), ("left_min", ctypes.c_uint64),
                ("max", ctypes.c_uint64), ("step", ctypes.c_uint64)]

    @classmethod
    def default(cls, **kw) -> "PairSchedule":
        """PairSchedule::default() (matching.rs:81-89)."""
        c = cls()
        load_library().skh_pair_schedule_default(ctypes.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    @classmethod
    def spec(cls) -> "PairSchedule":
        """PairSchedule::spec() (matching.rs:91-98)."""
        c = cls()
        load_library().skh_pair_schedule_spec(ctypes.byref(c))
        return c


class LengthMismatch(ValueError):
    """verified_matches' Err((matches.len(), inliers.len())) (tracks.rs:60-62)."""

    def __init__(self, n_matches: int, n_inliers: int):
        super().__init__((n_matches, n_inliers))
        self.lengths = (n_matches, n_inliers)


class ModelSelectionC(ctypes.Structure):
    _fields_ = [("has_selection", ctypes.c_int32), ("model", ctypes.c_int32),
                ("gric_f", ctypes.c_double), ("gric_h", ctypes.c_double),
                ("parallax", ctypes.c_uint64), ("parallax_needed", ctypes.c_uint64),
                ("sigma_px", ctypes.c_double)]


@dataclass
class ModelSelection:
    """matching.rs:380-394."""
    model: str  # "Fundamental" | "Homography"
    gric_f: float
    gric_h: float
    parallax: int
    parallax_needed: int
    sigma_px: float


@dataclass
class FundamentalFit:
    """matching.rs:911-930: F (3x3, unit Frobenius norm, rank 2), inlier flags,
    the model selection over the inliers, and the chosen model."""
    f: np.ndarray
    inliers: np.ndarray
    selection: Optional[ModelSelection]
    model: str

    def is_planar(self) -> bool:
        return self.model == "Homography"

_lib = None
_lib_lock = threading.Lock()


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads libskylens_hip.so and declares every exported signature."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ImportError(
                f"libskylens_hip.so not found at {path}: build it first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        # torch ships its own libamdhip64.so.7 under the same SONAME as
        # /opt/rocm's. Whichever loads first serves the whole process, and
        # torch's HIP stack fails to start on the other one, so torch (when
        # installed) is loaded first and this library binds to its runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(path)
        for name, res, args in EXPORTS:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


class SkhError(RuntimeError):
    pass


def _ptr(arr: np.ndarray) -> int:
    return arr.ctypes.data if arr.size else 0


class Context:
    """One skh_ctx: a HIP stream and grow-only device buffers on `device`."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _vp()
        rc = self.lib.skh_ctx_create(device, ctypes.byref(h))
        if rc != SKH_OK:
            raise SkhError(f"skh_ctx_create({device}) failed: status {rc} "
                           "(a gfx950 device is required)")
        self.handle = h
        self._plans = weakref.WeakSet()

    def close(self):
        """Destroys the context, after every plan still open on it (a plan's
        resources belong to its context, so it must go first)."""
        if self.handle:
            for plan in list(self._plans):
                plan.close()
            self.lib.skh_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # At interpreter shutdown the HIP runtime may already be gone; the
        # process exit releases everything. (The default argument keeps the
        # function reachable after module globals are cleared.)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc == SKH_OK:
            return
        msg = self.lib.skh_ctx_last_error(self.handle)
        msg = msg.decode() if msg else ""
        if rc == SKH_EINVAL:
            raise ValueError(f"{what}: {msg}")
        raise SkhError(f"{what}: status {rc}: {msg}")

    def set_tile(self, tile: int):
        """MFMA ranking tile (TILES index; -1 = automatic choice)."""
        self._check(self.lib.skh_ctx_set_tile(self.handle, int(tile)), "skh_ctx_set_tile")

    def autotune(self, precision: str = "bf16"):
        """Times the four ranking tiles per size class and loads the fastest:
        (table (25,) tile indices, ms (25, 4) per-launch times)."""
        table = np.zeros(TILE_CLASSES, np.int32)
        ms = np.zeros((TILE_CLASSES, 4), np.float64)
        self._check(self.lib.skh_ctx_autotune(self.handle, PRECISIONS[precision], _ptr(table),
                                              _ptr(ms)), "skh_ctx_autotune")
        return table, ms

    def set_tile_table(self, precision: str, table):
        t = None if table is None else np.ascontiguousarray(table, np.int32)
        self._check(self.lib.skh_ctx_set_tile_table(self.handle, PRECISIONS[precision],
                                                    None if t is None else _ptr(t)),
                    "skh_ctx_set_tile_table")

    def tile_table(self, precision: str = "bf16"):
        """(table (25,), tuned flag)."""
        table = np.zeros(TILE_CLASSES, np.int32)
        tuned = ctypes.c_int(0)
        self._check(self.lib.skh_ctx_tile_table(self.handle, PRECISIONS[precision], _ptr(table),
                                                ctypes.byref(tuned)), "skh_ctx_tile_table")
        return table, bool(tuned.value)

    @property
    def stream(self) -> int:
        return self.lib.skh_ctx_stream(self.handle) or 0

    # -- ratio_match ------------------------------------------------------
    def ratio_match(self, a, b, ratio: float, mutual: bool,
                    precision: str = "fp32") -> List[Tuple[int, int]]:
        ij = self.ratio_match_array(a, b, ratio, mutual, precision)
        return [(int(i), int(j)) for i, j in ij]

    def ratio_match_array(self, a, b, ratio: float, mutual: bool,
                          precision: str = "fp32") -> np.ndarray:
        a = _desc(a)
        b = _desc(b)
        if precision not in PRECISIONS:
            raise ValueError(f"unknown precision {precision!r}")
        cap = max(1, a.shape[0] if not mutual else min(a.shape[0], b.shape[0]))
        out = np.zeros((cap, 2), np.uint32)
        n = ctypes.c_size_t(0)
        rc = self.lib.skh_ratio_match(self.handle, _ptr(a), a.shape[0], _ptr(b), b.shape[0],
                                      ctypes.c_float(ratio), int(bool(mutual)),
                                      PRECISIONS[precision], _ptr(out), cap, ctypes.byref(n))
        self._check(rc, "skh_ratio_match")
        return out[: n.value].copy()

    def ratio_match_state(self, a, b):
        a = _desc(a)
        b = _desc(b)
        na, nb = a.shape[0], b.shape[0]
        j1 = np.zeros(na, np.uint32)
        d1 = np.zeros(na, np.float32)
        d2 = np.zeros(na, np.float32)
        ci = np.zeros(nb, np.uint32)
        cd = np.zeros(nb, np.float32)
        rc = self.lib.skh_ratio_match_state(self.handle, _ptr(a), na, _ptr(b), nb, _ptr(j1),
                                            _ptr(d1), _ptr(d2), _ptr(ci), _ptr(cd))
        self._check(rc, "skh_ratio_match_state")
        return j1, d1, d2, ci, cd

    # -- ransac_fundamental ----------------------------------------------
    def ransac_fundamental(self, x1, x2, cfg: Optional[RansacConfig] = None
                           ) -> Optional[FundamentalFit]:
        x1 = np.ascontiguousarray(np.asarray(x1, dtype=np.float64).reshape(-1, 2))
        x2 = np.ascontiguousarray(np.asarray(x2, dtype=np.float64).reshape(-1, 2))
        if x1.shape[0] != x2.shape[0]:
            return None  # matching.rs:947: length mismatch -> None
        cfg = cfg or RansacConfig.default()
        n = x1.shape[0]
        F = np.zeros(9, np.float64)
        inl = np.zeros(max(n, 1), np.uint8)
        sel = ModelSelectionC()
        found = ctypes.c_int(0)
        rc = self.lib.skh_ransac_fundamental(self.handle, _ptr(x1), _ptr(x2), n, ctypes.byref(cfg),
                                             _ptr(F), _ptr(inl), ctypes.byref(sel),
                                             ctypes.byref(found))
        self._check(rc, "skh_ransac_fundamental")
        if not found.value:
            return None
        ms = None
        if sel.has_selection:
            ms = ModelSelection("Homography" if sel.model else "Fundamental", sel.gric_f,
                                sel.gric_h, int(sel.parallax), int(sel.parallax_needed),
                                sel.sigma_px)
        return FundamentalFit(F.reshape(3, 3), inl[:n].astype(bool), ms,
                              ms.model if ms else "Fundamental")

    @property
    def last_ransac_iters(self) -> int:
        return int(self.lib.skh_ctx_last_ransac_iters(self.handle))

    def ratio_match_dev(self, a_ptr: int, na: int, b_ptr: int, nb: int, ratio: float,
                        mutual: bool, precision: str, out_ij_ptr: int, out_len_ptr: int):
        """Device-pointer form: inputs/outputs already in HBM, async on `stream`."""
        rc = self.lib.skh_ratio_match_dev(self.handle, a_ptr, na, b_ptr, nb, ctypes.c_float(ratio),
                                          int(bool(mutual)), PRECISIONS[precision], out_ij_ptr,
                                          out_len_ptr)
        self._check(rc, "skh_ratio_match_dev")


    def ratio_match_dev_noalloc(self, a_ptr: int, na: int, b_ptr: int, nb: int, ratio: float,
                                mutual: bool, precision: str, out_ij_ptr: int, out_len_ptr: int):
        """ratio_match_dev without allocating (fixed per-GPU scratch; at most 8192
        rows per side; calls overlapping on one GPU race)."""
        rc = self.lib.skh_ratio_match_dev_noalloc(self.handle, a_ptr, na, b_ptr, nb,
                                                  ctypes.c_float(ratio), int(bool(mutual)),
                                                  PRECISIONS[precision], out_ij_ptr, out_len_ptr)
        self._check(rc, "skh_ratio_match_dev_noalloc")


    def detect_and_describe(self, img, cfg: Optional[DetectorConfig] = None,
                            describe: bool = True):
        """detect_and_describe (features.rs:594-648) of a (h, w) float32 image:
        (keypoints (n, 5) float32 in KEYPOINT_FIELDS order, descriptors
        (n, 128) float32 or None). ``describe=False`` is detect()
        (features.rs:487-493)."""
        img = np.ascontiguousarray(img, dtype=np.float32)
        if img.ndim != 2:
            raise ValueError("image must be a (height, width) array")
        cfg = cfg or DetectorConfig.default()
        h, w = img.shape
        cap = max(int(cfg.max_features), 1)
        kps = np.empty((cap, 5), np.float32)
        desc = np.empty((cap, 128), np.float32) if describe else None
        n = ctypes.c_size_t(0)
        self._check(self.lib.skh_detect_and_describe(
            self.handle, _ptr(img), w, h, ctypes.byref(cfg), _ptr(kps),
            _ptr(desc) if describe else None, cap, ctypes.byref(n)), "skh_detect_and_describe")
        k = n.value
        return kps[:k].copy(), (desc[:k].copy() if describe else None)

    def build_tracks(self, pairs, keypoints, cfg: Optional[TrackConfig] = None):
        """build_tracks (tracks.rs:190-508). ``pairs``: (image_a, image_b,
        [(fa, fb), ...]) per PairMatches; ``keypoints``: one (n_i, 2) array of
        pixels per image. Returns (tracks, stats): each track a list of
        (image, feature) observations (the pixel is keypoints[image][feature]),
        stats a dict of TrackStats."""
        packed = TrackInput(pairs, keypoints)
        nt, st = self.build_tracks_packed(packed, cfg)
        toff, obs = packed.track_off, packed.obs
        tracks = [[(int(a), int(b)) for a, b in obs[toff[t]:toff[t + 1]]] for t in range(nt)]
        return tracks, st

    def build_tracks_packed(self, packed: "TrackInput", cfg: Optional[TrackConfig] = None):
        """skh_build_tracks on pre-packed inputs: (track count, stats dict); the
        tracks are left in packed.track_off / packed.obs."""
        cfg = cfg or TrackConfig.default()
        q = packed
        nt = ctypes.c_size_t(0)
        st = TrackStats()
        self._check(self.lib.skh_build_tracks(
            self.handle, q.n_pairs, _ptr(q.ia), _ptr(q.ib), _ptr(q.m_off), _ptr(q.m),
            q.n_images, _ptr(q.kp_off), _ptr(q.kp), ctypes.byref(cfg), _ptr(q.track_off),
            q.track_off.shape[0], _ptr(q.obs), q.n, ctypes.byref(nt), ctypes.byref(st)),
            "skh_build_tracks")
        return nt.value, {k: int(getattr(st, k)) for k in TRACK_STATS}

    def detect_and_describe_dev(self, img_ptr: int, w: int, h: int, cfg: DetectorConfig,
                                kp_ptr: int, desc_ptr: int, cap: int) -> int:
        """Device-resident form: returns the feature count."""
        n = ctypes.c_size_t(0)
        self._check(self.lib.skh_detect_and_describe_dev(
            self.handle, img_ptr, w, h, ctypes.byref(cfg), kp_ptr, desc_ptr or None, cap,
            ctypes.byref(n)), "skh_detect_and_describe_dev")
        return n.value


class BatchDesc(ctypes.Structure):
    """skh_batch_desc."""
    _fields_ = [("n_views", ctypes.c_uint32), ("view_offset", _vp), ("d_desc", _vp),
                ("d_kp", _vp), ("n_pairs", ctypes.c_uint32), ("pairs", _vp),
                ("ratio", ctypes.c_float), ("mutual", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("run_ransac", ctypes.c_int32),
                ("px_offset", ctypes.c_double), ("ransac", RansacConfig)]


@dataclass
class BatchResult:
    match_base: np.ndarray   # (P + 1,) uint64 capacity offsets
    match_len: np.ndarray    # (P,) uint32
    match_ij: np.ndarray     # (cap, 2) uint32
    inliers: Optional[np.ndarray]
    F: Optional[np.ndarray]  # (P, 3, 3)
    found: Optional[np.ndarray]
    sel: Optional[np.ndarray]

    def matches(self, p: int) -> np.ndarray:
        b = int(self.match_base[p])
        return self.match_ij[b:b + int(self.match_len[p])]

    def inlier_mask(self, p: int) -> np.ndarray:
        b = int(self.match_base[p])
        return self.inliers[b:b + int(self.match_len[p])].astype(bool)


class Plan:
    """A batched matching + RANSAC plan over device-resident views (skh_plan).

    `desc_ptr`/`kp_ptr` are device addresses (e.g. torch tensor data_ptr())
    that must outlive the plan; `view_offset` and `pairs` are host arrays.
    """

    def __init__(self, ctx: Context, view_offset, desc_ptr: int, kp_ptr: int, pairs,
                 ratio: float = 0.8, mutual: bool = True, precision: str = "bf16",
                 run_ransac: bool = True, px_offset: float = 0.0,
                 cfg: Optional[RansacConfig] = None):
        self.ctx = ctx
        self.lib = ctx.lib
        self._off = np.ascontiguousarray(view_offset, dtype=np.uint64)
        self._pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1, 2))
        d = BatchDesc()
        d.n_views = len(self._off) - 1
        d.view_offset = self._off.ctypes.data
        d.d_desc = desc_ptr
        d.d_kp = kp_ptr
        d.n_pairs = self._pairs.shape[0]
        d.pairs = _ptr(self._pairs)
        d.ratio = ratio
        d.mutual = int(bool(mutual))
        d.precision = PRECISIONS[precision]
        d.run_ransac = int(bool(run_ransac))
        d.px_offset = px_offset
        d.ransac = cfg or RansacConfig.default()
        self.n_pairs = d.n_pairs
        self.run_ransac = bool(run_ransac)
        h = _vp()
        rc = self.lib.skh_plan_create(ctx.handle, ctypes.byref(d), ctypes.byref(h))
        ctx._check(rc, "skh_plan_create")
        self.handle = h
        ctx._plans.add(self)
        self.match_base = np.zeros(self.n_pairs + 1, np.uint64)
        ctx._check(self.lib.skh_plan_match_base(self.handle, _ptr(self.match_base)),
                   "skh_plan_match_base")

    def run(self):
        self.ctx._check(self.lib.skh_plan_run(self.handle), "skh_plan_run")

    def sync(self):
        self.ctx._check(self.lib.skh_plan_sync(self.handle), "skh_plan_sync")

    def set_timing(self, on: bool):
        self.ctx._check(self.lib.skh_plan_set_timing(self.handle, int(on)), "skh_plan_set_timing")

    def stage_ms(self, stage: int) -> float:
        return float(self.lib.skh_plan_stage_ms(self.handle, stage))

    def ransac_profile(self):
        """(iters per pair, phase ticks (P, 11), ticks per ms) of the last run."""
        it = np.zeros(max(self.n_pairs, 1), np.uint64)
        ph = np.zeros((max(self.n_pairs, 1), 11), np.uint64)
        tpm = ctypes.c_double(0.0)
        self.ctx._check(self.lib.skh_plan_ransac_profile(self.handle, _ptr(it), _ptr(ph),
                                                         ctypes.byref(tpm)), "ransac_profile")
        return it[: self.n_pairs], ph[: self.n_pairs], tpm.value

    def verified(self):
        """(verified_len (P,), verified_ij (cap, 2)) of the last run: each pair's
        inlier matches, in match order, at its match_base offset
        (skh_plan_verified; tracks.rs:56-69 per pair)."""
        P = self.n_pairs
        cap = int(self.match_base[-1])
        vl = np.zeros(max(P, 1), np.uint32)
        vij = np.zeros((max(cap, 1), 2), np.uint32)
        self.ctx._check(self.lib.skh_plan_verified(self.handle, _ptr(vl), _ptr(vij)),
                        "skh_plan_verified")
        return vl[:P], vij

    def tiles(self) -> dict:
        """Pairs ranked with each tile ({"256x64": count, ...})."""
        n = np.zeros(4, np.uint32)
        self.ctx._check(self.lib.skh_plan_tiles(self.handle, _ptr(n)), "skh_plan_tiles")
        return {TILE_NAMES[k]: int(n[k]) for k in range(4) if n[k]}

    def lowp_state(self) -> dict:
        """The BF16/FP8 ranking state of the last run (skh_plan_lowp_state):
        per-row keys k1, k2 (~ d^2 / 2) and j1, per-column (c1, c2), per-pair
        error bound eps, and the number of rows the exact re-score took."""
        R = int(np.sum(self._rows()))
        C = int(np.sum(self._cols()))
        k1 = np.zeros(max(R, 1), np.float32)
        k2 = np.zeros(max(R, 1), np.float32)
        j1 = np.zeros(max(R, 1), np.uint32)
        ct = np.zeros((max(C, 1), 2), np.float32)
        eps = np.zeros(max(self.n_pairs, 1), np.float64)
        q = ctypes.c_uint64(0)
        self.ctx._check(self.lib.skh_plan_lowp_state(self.handle, _ptr(k1), _ptr(j1), _ptr(k2),
                                                     _ptr(ct), _ptr(eps), ctypes.byref(q)),
                        "skh_plan_lowp_state")
        rb = np.concatenate([[0], np.cumsum(self._rows())])
        cb = np.concatenate([[0], np.cumsum(self._cols())])
        return {"k1": k1[:R], "k2": k2[:R], "j1": j1[:R], "col": ct[:C], "eps": eps[:self.n_pairs],
                "queued": int(q.value), "row_base": rb, "col_base": cb}

    def _rows(self):
        return np.array([self._off[i + 1] - self._off[i] for i, _ in self._pairs], np.int64)

    def _cols(self):
        return np.array([self._off[j + 1] - self._off[j] for _, j in self._pairs], np.int64)

    def time_distance(self, reps: int = 3) -> float:
        """Average ms of the distance kernel alone (HIP events, plan stream)."""
        return float(self.lib.skh_plan_time_distance(self.handle, int(reps)))

    def results(self) -> BatchResult:
        P = self.n_pairs
        cap = int(self.match_base[-1])
        ml = np.zeros(max(P, 1), np.uint32)
        ij = np.zeros((max(cap, 1), 2), np.uint32)
        inl = F = found = sel = None
        if self.run_ransac:
            inl = np.zeros(max(cap, 1), np.uint8)
            F = np.zeros((max(P, 1), 9), np.float64)
            found = np.zeros(max(P, 1), np.int32)
            sel = (ModelSelectionC * max(P, 1))()
        rc = self.lib.skh_plan_results(
            self.handle, _ptr(ml), _ptr(ij), _ptr(inl) if inl is not None else None,
            _ptr(F) if F is not None else None, _ptr(found) if found is not None else None,
            ctypes.addressof(sel) if sel is not None else None)
        self.ctx._check(rc, "skh_plan_results")
        return BatchResult(self.match_base, ml[:P], ij, inl, None if F is None else F[:P].reshape(P, 3, 3),
                           None if found is None else found[:P], sel)

    def close(self):
        if self.handle:
            if self.ctx.handle:
                self.lib.skh_plan_destroy(self.handle)
            self.handle = None

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def _desc(x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    if x.ndim == 1 and x.size == 0:
        x = x.reshape(0, 128)
    if x.ndim != 2 or x.shape[1] != 128:
        raise ValueError(f"descriptors must be (n, 128) float32, got {x.shape}")
    return x


_default = None
_default_lock = threading.Lock()


def default_context() -> Context:
    global _default
    with _default_lock:
        if _default is None:
            _default = Context(0)
        return _default


def available_precisions() -> List[str]:
    """Precisions this build of the library implements."""
    lib = load_library()
    return [k for k, v in PRECISIONS.items() if lib.skh_precision_supported(v)]


def candidate_pairs(views: Sequence[Tuple[int, int]], temporal: int, cross: int,
                    pow2_max: int) -> List[Tuple[int, int]]:
    """matching.rs:18-40. `views[k] = (camera, position)`; pairs (i, j), i < j, sorted."""
    lib = load_library()
    v = np.ascontiguousarray(np.asarray(views, dtype=np.uint64).reshape(-1, 2))
    n = ctypes.c_size_t(0)
    mx = 0xFFFFFFFFFFFFFFFF
    rc = lib.skh_candidate_pairs(_ptr(v), v.shape[0], min(temporal, mx), min(cross, mx),
                                 min(pow2_max, mx), None, 0, ctypes.byref(n))
    if rc not in (SKH_OK, SKH_EINVAL):
        raise SkhError(f"skh_candidate_pairs: {rc}")
    out = np.zeros((max(1, n.value), 2), np.uint64)
    rc = lib.skh_candidate_pairs(_ptr(v), v.shape[0], min(temporal, mx), min(cross, mx),
                                 min(pow2_max, mx), _ptr(out), n.value, ctypes.byref(n))
    if rc != SKH_OK:
        raise SkhError(f"skh_candidate_pairs: {rc}")
    return [(int(i), int(j)) for i, j in out[: n.value]]


def scheduled_pairs(views: Sequence[Tuple[int, int]],
                    sch: Optional[PairSchedule] = None) -> List[Tuple[int, int]]:
    """matching.rs:103-142: the production pair rule; PairSchedule::default() if
    `sch` is None."""
    lib = load_library()
    sch = sch or PairSchedule.default()
    v = np.ascontiguousarray(np.asarray(views, dtype=np.uint64).reshape(-1, 2))
    n = ctypes.c_size_t(0)
    rc = lib.skh_scheduled_pairs(_ptr(v), v.shape[0], ctypes.byref(sch), None, 0, ctypes.byref(n))
    if rc not in (SKH_OK, SKH_EINVAL):
        raise SkhError(f"skh_scheduled_pairs: {rc}")
    out = np.zeros((max(1, n.value), 2), np.uint64)
    rc = lib.skh_scheduled_pairs(_ptr(v), v.shape[0], ctypes.byref(sch), _ptr(out), n.value,
                                 ctypes.byref(n))
    if rc != SKH_OK:
        raise SkhError(f"skh_scheduled_pairs: {rc}")
    return [(int(i), int(j)) for i, j in out[: n.value]]


def verified_matches(matches, inliers) -> List[Tuple[int, int]]:
    """tracks.rs:56-69: the matches whose inlier flag is set; raises
    LengthMismatch((len(matches), len(inliers))) for the reference's Err."""
    lib = load_library()
    m = np.ascontiguousarray(np.asarray(matches, dtype=np.uint32).reshape(-1, 2))
    f = np.ascontiguousarray(np.asarray(inliers, dtype=np.uint8).reshape(-1))
    if m.shape[0] != f.shape[0]:
        raise LengthMismatch(m.shape[0], f.shape[0])
    out = np.zeros((max(1, m.shape[0]), 2), np.uint32)
    n = ctypes.c_size_t(0)
    rc = lib.skh_verified_matches(_ptr(m), m.shape[0], _ptr(f), f.shape[0], _ptr(out),
                                  ctypes.byref(n))
    if rc != SKH_OK:
        raise SkhError(f"skh_verified_matches: {rc}")
    return [(int(i), int(j)) for i, j in out[: n.value]]


def ratio_match(a, b, ratio: float, mutual: bool, precision: str = "fp32") -> List[Tuple[int, int]]:
    """matching.rs:167-215: (a index, b index) in a-index order."""
    return default_context().ratio_match(a, b, ratio, mutual, precision)


def ransac_fundamental(x1, x2, cfg: Optional[RansacConfig] = None) -> Optional[FundamentalFit]:
    """matching.rs:941-1104: Some(FundamentalFit) or None."""
    return default_context().ransac_fundamental(x1, x2, cfg)
