"""FP64 check of the low-precision ranking's error bound (numpy only).

For a BF16 or FP8 plan, skh_plan_lowp_state returns every row's two smallest
keys (key ~ d^2 / 2), the column of the smallest, every column's two smallest
keys, and the pair's error bound eps, which must satisfy |2 key - d^2| <= eps
for every element of the pair. Here d^2 is recomputed in float64 from the f32
descriptors (|a|^2 + |b|^2 - 2 a.b), and the order statistics it implies are
compared with the keys: the row minimum and its column, the row's second
smallest, and each column's two smallest (the squared distance of
crates/core/src/matching.rs:144-155, in exact arithmetic).
"""
from __future__ import annotations

import numpy as np


def _two_smallest(d: np.ndarray, axis: int):
    if d.shape[axis] < 2:
        first = d.min(axis=axis)
        return first, np.full_like(first, np.inf)
    part = np.partition(d, 1, axis=axis)
    return np.take(part, 0, axis=axis), np.take(part, 1, axis=axis)


def pair_errors(a: np.ndarray, b: np.ndarray, st: dict, p: int) -> dict:
    """Largest |2 key - d^2_fp64| over pair p's row and column top-2 keys, and
    its ratio to the pair's eps."""
    a64 = a.astype(np.float64)
    b64 = b.astype(np.float64)
    d = (np.einsum("ij,ij->i", a64, a64)[:, None] + np.einsum("ij,ij->i", b64, b64)[None, :]
         - 2.0 * (a64 @ b64.T))
    r0, c0 = int(st["row_base"][p]), int(st["col_base"][p])
    na, nb = a.shape[0], b.shape[0]
    k1 = 2.0 * st["k1"][r0:r0 + na].astype(np.float64)
    k2 = 2.0 * st["k2"][r0:r0 + na].astype(np.float64)
    j1 = st["j1"][r0:r0 + na].astype(np.int64)
    c = 2.0 * st["col"][c0:c0 + nb].astype(np.float64)
    d1, d2 = _two_smallest(d, 1)
    e1, e2 = _two_smallest(d, 0)
    errs = [np.abs(k1 - d1), np.abs(c[:, 0] - e1)]
    if nb >= 2:
        errs.append(np.abs(k2 - d2))
        errs.append(np.abs(k1 - d[np.arange(na), np.minimum(j1, nb - 1)]))
    if na >= 2:
        errs.append(np.abs(c[:, 1] - e2))
    worst = max(float(np.max(e)) for e in errs if e.size)
    eps = float(st["eps"][p])
    return {"max_err": worst, "eps": eps, "ratio": worst / eps if np.isfinite(eps) else 0.0,
            "j1_ok": bool(nb < 2 or np.all(d[np.arange(na), j1.clip(0, nb - 1)] <= d1 + eps))}


def scene_errors(scene, st: dict, pairs) -> dict:
    """pair_errors over the listed pair indices of a workloads.Scene."""
    rs = [pair_errors(scene.view_desc(int(scene.pairs[p][0])),
                      scene.view_desc(int(scene.pairs[p][1])), st, int(p)) for p in pairs]
    finite = [r for r in rs if np.isfinite(r["eps"])]
    return {"pairs_checked": len(rs),
            "max_d2_err_fp64": max((r["max_err"] for r in finite), default=0.0),
            "max_err_over_eps": max((r["ratio"] for r in finite), default=0.0),
            "eps_max": float(np.max(st["eps"][np.isfinite(st["eps"])], initial=0.0)),
            "eps_median": float(np.median(st["eps"][np.isfinite(st["eps"])]))
            if np.any(np.isfinite(st["eps"])) else 0.0,
            "j1_within_eps": all(r["j1_ok"] for r in rs)}
