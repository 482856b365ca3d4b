"""Writes the golden fixtures of the matching path from the oracle.

The oracle is the C restatement of crates/core/src/matching.rs; the
fixtures freeze its outputs on seeded inputs built with the reference's own
test generators (matching.rs:1141-1220, :1580-1656) so that a later change of
either the oracle or the HIP path shows up as a fixture mismatch:

* ``matches.npz``  -- ratio_match(match_fixture(seed, na, extra), ratio,
  mutual) for a few (seed, na, extra, ratio, mutual) cases: the (i, j) lists;
* ``ransac.npz``   -- ransac_fundamental on correspondences / planar /
  ground-with-buildings / line-plus-general sets: found flag, F bits
  (uint64 view of the row-major doubles), inlier masks, hypotheses drawn.

Run from the repository root: ``python tests/golden/make_golden.py``.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind as ob  # noqa: E402

MATCH_CASES = [(21, 700, 300, 0.8, True), (5, 300, 100, 0.95, False), (93, 257, 64, 0.6, True),
               (7, 129, 40, 1.0, True)]


def ransac_inputs():
    """(label, x1, x2) for each RANSAC fixture case."""
    out = []
    for n, sigma, frac, seed in [(300, 0.5, 0.3, 5), (300, 0.5, 0.5, 9), (120, 0.5, 0.2, 17),
                                 (600, 0.5, 0.4, 23)]:
        x1, x2, _, _, _ = ob.correspondences(n, sigma, frac, seed)
        out.append((f"corr_{n}_{frac}_{seed}", x1, x2))
    x1, x2 = ob.planar_correspondences(200, 0.5, 31)[:2]
    out.append(("planar_200_31", x1, x2))
    x1, x2 = ob.ground_with_buildings(240, 3, 0.5, 37)[:2]
    out.append(("buildings_240_37", x1, x2))
    x1, x2 = ob.line_plus_general(40, 0.5, 41)[:2]
    out.append(("line_plus_general_40_41", x1, x2))
    return out


def main():
    ob.build_oracle()
    m = {}
    for k, (seed, na, extra, ratio, mutual) in enumerate(MATCH_CASES):
        a, b = ob.match_fixture(seed, na, extra)
        ij = np.asarray(ob.ratio_match(a, b, ratio, mutual), np.uint32).reshape(-1, 2)
        m[f"case{k}"] = ij
        m[f"case{k}_args"] = np.array([seed, na, extra, ratio, int(mutual)], np.float64)
    np.savez_compressed(os.path.join(HERE, "matches.npz"), **m)

    r = {}
    for label, x1, x2 in ransac_inputs():
        res = ob.ransac_fundamental(x1, x2)
        st = ob.ransac_last_stats()
        r[f"{label}_x1"] = x1
        r[f"{label}_x2"] = x2
        r[f"{label}_found"] = np.array([res is not None], np.bool_)
        r[f"{label}_iters"] = np.array([st["iters"]], np.uint64)
        if res is not None:
            F, inl, _ = res
            r[f"{label}_F"] = np.ascontiguousarray(F.reshape(9)).view(np.uint64)
            r[f"{label}_inliers"] = inl.astype(np.bool_)
    np.savez_compressed(os.path.join(HERE, "ransac.npz"), **r)
    print("wrote", sorted(m)[:2], "...", len(r), "ransac arrays")


if __name__ == "__main__":
    main()
