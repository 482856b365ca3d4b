"""Writes the per-pair golden fixtures of the full-size workloads from the oracle.

For the single pair (800 x 800 descriptors with keypoints tied to them), the
full scene (240 views, 1500-4096 features per view, 3663 candidate pairs) and
the uneven batch (264 views, 200-8000 features per view, 4000 pairs), the
oracle -- the C restatement of crates/core/src/matching.rs -- runs every pair
the way benches/pipeline.rs:247-295 does: ratio_match(view i, view j, 0.8,
mutual), the keypoint hand-off (f32 -> f64, no +0.5), then
ransac_fundamental with RansacConfig::default(). Per pair the fixture holds:

* ``match_count`` and ``match_hash``: the length of the (i, j) list and the
  reference's FNV-1a hash of it (matching.rs:1867-1878);
* ``found`` (1 = Some, 0 = None), ``iters`` (hypotheses drawn), ``n_inliers``,
  ``model`` (-1 no selection, 0 Fundamental, 1 Homography);
* ``f_hash`` / ``mask_hash``: 64-bit BLAKE2b digests of the row-major F bits
  and of the inlier mask bytes (0 for None);
* ``input_hash``: digests of the generated descriptor, keypoint, offset and
  pair arrays, so a generator drift on another machine shows up as such and
  not as a parity failure.

Run from the repository root: ``python tests/golden/make_scene_golden.py``
(about two minutes on 8 cores).
"""
import hashlib
import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "skylens-stream-rs_amd"))
import oracle_bind as ob  # noqa: E402
import workloads  # noqa: E402


def digest(*arrays) -> int:
    h = hashlib.blake2b(digest_size=8)
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return int.from_bytes(h.digest(), "little")


def scene_inputs_hash(sc) -> np.ndarray:
    return np.array([digest(sc.desc), digest(sc.kp), digest(np.asarray(sc.offsets, np.int64)),
                     digest(np.asarray(sc.pairs, np.int64))], np.uint64)


def pair_record(a, b, kp_a, kp_b):
    """(match_count, match_hash, found, iters, n_inliers, model, f_hash, mask_hash)."""
    m = ob.ratio_match(a, b, 0.8, True)
    mm = np.asarray(m, np.int64).reshape(-1, 2)
    x1 = kp_a[mm[:, 0]].astype(np.float64)
    x2 = kp_b[mm[:, 1]].astype(np.float64)
    r = ob.ransac_fundamental(x1, x2)
    iters = ob.ransac_last_stats()["iters"]
    if r is None:
        return (len(m), ob.pairs_hash(m), 0, iters, 0, -1, 0, 0)
    F, inl, sel = r
    model = (1 if sel.model else 0) if sel.has_selection else -1
    return (len(m), ob.pairs_hash(m), 1, iters, int(inl.sum()), model,
            digest(np.ascontiguousarray(F.reshape(9))), digest(inl.astype(np.uint8)))


_SC = None


def _scene_pair(p):
    i, j = (int(v) for v in _SC.pairs[p])
    return pair_record(_SC.view_desc(i), _SC.view_desc(j), _SC.view_kp(i), _SC.view_kp(j))


FIELDS = [("match_count", np.uint32), ("match_hash", np.uint64), ("found", np.int8),
          ("iters", np.uint32), ("n_inliers", np.uint32), ("model", np.int8),
          ("f_hash", np.uint64), ("mask_hash", np.uint64)]


def to_arrays(records):
    cols = list(zip(*records))
    return {name: np.asarray(c, dt) for (name, dt), c in zip(FIELDS, cols)}


def main():
    global _SC
    ob.build_oracle()
    sp = workloads.single_pair(0)
    rec = pair_record(sp.a, sp.b, sp.kp_a, sp.kp_b)
    np.savez_compressed(os.path.join(HERE, "single_pair.npz"),
                        input_hash=np.array([digest(sp.a), digest(sp.b), digest(sp.kp_a),
                                             digest(sp.kp_b)], np.uint64),
                        **to_arrays([rec]))
    print("single pair:", rec, flush=True)
    for name, make in (("full", lambda: workloads.full_scene(seed=1)),
                       ("uneven", lambda: workloads.uneven_scene(seed=4))):
        _SC = make()
        P = len(_SC.pairs)
        with mp.get_context("fork").Pool(min(8, os.cpu_count() or 1)) as pool:
            recs = pool.map(_scene_pair, range(P), chunksize=8)
        out = to_arrays(recs)
        np.savez_compressed(os.path.join(HERE, f"scene_{name}.npz"),
                            input_hash=scene_inputs_hash(_SC), **out)
        print(f"{name}: {P} pairs, {int(out['match_count'].sum())} matches, "
              f"{int(out['found'].sum())} verified", flush=True)


if __name__ == "__main__":
    main()
