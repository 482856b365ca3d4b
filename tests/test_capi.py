"""C-ABI checks that need no GPU: the library loads, exports every symbol the
headers declare, and the host-only entry points match the oracle."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import oracle_bind as ob
import skylens_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(skh_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = skylens_hip.load_library()
    declared = declared_symbols()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in skylens_hip.EXPORTS}
    assert declared == bound, (declared ^ bound)


def _grid(npos, ncam=3):
    return [(c, p) for p in range(npos) for c in range(ncam)]


@pytest.mark.parametrize("args", [(10, 2, 1, 0), (80, 5, 4, 16), (80, 5, 4, 2**64 - 1),
                                  (80, 5, 4, 0), (7, 0, 0, 0), (0, 5, 4, 16), (1, 5, 4, 16)])
def test_candidate_pairs_matches_oracle(args):
    npos, t, c, p = args
    views = _grid(npos)
    assert skylens_hip.candidate_pairs(views, t, c, p) == ob.candidate_pairs(views, t, c, p)


def test_candidate_pairs_unsorted_views():
    rng = np.random.default_rng(3)
    views = [(int(rng.integers(0, 3)), int(rng.integers(0, 40))) for _ in range(150)]
    assert skylens_hip.candidate_pairs(views, 5, 4, 16) == ob.candidate_pairs(views, 5, 4, 16)


def test_candidate_pairs_capacity_error():
    lib = skylens_hip.load_library()
    v = np.array(_grid(10), np.uint64)
    out = np.zeros((3, 2), np.uint64)
    n = ctypes.c_size_t(0)
    rc = lib.skh_candidate_pairs(v.ctypes.data, v.shape[0], 2, 1, 0, out.ctypes.data, 3,
                                 ctypes.byref(n))
    assert rc == skylens_hip.SKH_EINVAL
    assert n.value == 135


def test_null_context_is_rejected():
    lib = skylens_hip.load_library()
    n = ctypes.c_size_t(0)
    assert lib.skh_ratio_match(None, None, 0, None, 0, ctypes.c_float(0.8), 1, 0, None, 0,
                               ctypes.byref(n)) == skylens_hip.SKH_EINVAL


def _c_layout():
    """sizeof / offsetof of the header's structs, as a C compiler lays them out."""
    import subprocess
    import tempfile

    fields = {
        "skh_ransac_cfg": ["threshold_px", "min_inlier_ratio", "max_iters", "confidence", "seed",
                           "min_inliers", "max_chance_prob"],
        "skh_model_sel": ["has_selection", "model", "gric_f", "gric_h", "parallax",
                          "parallax_needed", "sigma_px"],
        "skh_batch_desc": ["n_views", "view_offset", "d_desc", "d_kp", "n_pairs", "pairs", "ratio",
                           "mutual", "precision", "run_ransac", "px_offset", "ransac"],
        "skh_pair_schedule": ["temporal", "pow2_max", "cross_kind", "range", "right_min",
                              "left_min", "max", "step"],
        "skh_detector_cfg": ["octaves", "scales", "sigma0", "contrast", "edge_ratio",
                             "upscale_fill", "max_features", "upscale_below"],
        "skh_keypoint": ["x", "y", "sigma", "response", "angle"],
        "skh_track_cfg": ["policy", "pad", "min_length", "max_tracks"],
        "skh_track_stats": list(skylens_hip.TRACK_STATS),
    }
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "skylens_hip.h"',
             "int main(void) {"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        with open(src, "w") as f:
            f.write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (l.split() for l in out.splitlines())}, fields


def test_struct_layout_matches_c_header():
    """The ctypes mirrors (and so the Rust #[repr(C)] binding written from the same
    header) lay out every field where the C compiler does."""
    got, fields = _c_layout()
    mirrors = {"skh_ransac_cfg": skylens_hip.RansacConfig,
               "skh_model_sel": skylens_hip.ModelSelectionC,
               "skh_batch_desc": skylens_hip.BatchDesc,
               "skh_pair_schedule": skylens_hip.PairSchedule,
               "skh_detector_cfg": skylens_hip.DetectorConfig,
               "skh_track_cfg": skylens_hip.TrackConfig,
               "skh_track_stats": skylens_hip.TrackStats}
    assert got["skh_keypoint"] == 4 * len(skylens_hip.KEYPOINT_FIELDS)
    for i, f in enumerate(skylens_hip.KEYPOINT_FIELDS):
        assert got[f"skh_keypoint.{f}"] == 4 * i
    for st, cls in mirrors.items():
        assert ctypes.sizeof(cls) == got[st], st
        for f in fields[st]:
            assert getattr(cls, f).offset == got[f"{st}.{f}"], f"{st}.{f}"
