"""Helpers shared by the CPU and GPU tests of the full-size per-pair fixtures
(tests/golden/scene_*.npz, single_pair.npz, written by
tests/golden/make_scene_golden.py from the oracle)."""
import hashlib
import os

import numpy as np

import oracle_bind as ob

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def digest(*arrays) -> int:
    h = hashlib.blake2b(digest_size=8)
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return int.from_bytes(h.digest(), "little")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def scene(name):
    import workloads

    return {"full": lambda: workloads.full_scene(seed=1),
            "uneven": lambda: workloads.uneven_scene(seed=4)}[name]()


def scene_inputs_hash(sc) -> np.ndarray:
    return np.array([digest(sc.desc), digest(sc.kp), digest(np.asarray(sc.offsets, np.int64)),
                     digest(np.asarray(sc.pairs, np.int64))], np.uint64)


def plan_mismatches(res, z, iters=None):
    """Pairs whose plan result differs from the fixture: (pair, what) list."""
    bad = []
    P = len(z["match_count"])
    for p in range(P):
        m = res.matches(p)
        if len(m) != int(z["match_count"][p]):
            bad.append((p, f"match count {len(m)} != {int(z['match_count'][p])}"))
            continue
        if ob.pairs_hash(np.asarray(m, np.uint64)) != int(z["match_hash"][p]):
            bad.append((p, "match list"))
            continue
        if res.found is None:
            continue
        if int(res.found[p]) != int(z["found"][p]):
            bad.append((p, f"found {int(res.found[p])} != {int(z['found'][p])}"))
            continue
        if iters is not None and int(iters[p]) != int(z["iters"][p]):
            bad.append((p, f"iterations {int(iters[p])} != {int(z['iters'][p])}"))
        if int(z["found"][p]) == 1:
            if digest(np.ascontiguousarray(res.F[p].reshape(9))) != int(z["f_hash"][p]):
                bad.append((p, "F bits"))
            if digest(res.inlier_mask(p).astype(np.uint8)) != int(z["mask_hash"][p]):
                bad.append((p, "inlier mask"))
    return bad
