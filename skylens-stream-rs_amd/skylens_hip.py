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
import threading
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
    ("skh_ratio_match_state", ctypes.c_int,
     [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, _vp, _vp, _vp, _vp, _vp]),
]

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

    def close(self):
        if self.handle:
            self.lib.skh_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
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

    def ratio_match_dev(self, a_ptr: int, na: int, b_ptr: int, nb: int, ratio: float,
                        mutual: bool, precision: str, out_ij_ptr: int, out_len_ptr: int):
        """Device-pointer form: inputs/outputs already in HBM, async on `stream`."""
        rc = self.lib.skh_ratio_match_dev(self.handle, a_ptr, na, b_ptr, nb, ctypes.c_float(ratio),
                                          int(bool(mutual)), PRECISIONS[precision], out_ij_ptr,
                                          out_len_ptr)
        self._check(rc, "skh_ratio_match_dev")


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


def ratio_match(a, b, ratio: float, mutual: bool, precision: str = "fp32") -> List[Tuple[int, int]]:
    """matching.rs:167-215: (a index, b index) in a-index order."""
    return default_context().ratio_match(a, b, ratio, mutual, precision)
