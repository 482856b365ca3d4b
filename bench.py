"""Matched image pairs/sec of the pairwise matching path on one MI355X.

Headline workload: the uneven batch -- 4000 image pairs of a 264-view scene
(3 cameras x 88 positions, candidate_pairs(5, 4, 16), the first 4000 pairs),
per-view feature counts log-uniform in [200, 8000], unit 128-d descriptors
(60 % noisy copies of a 20k pool), pair inlier ratios 10-90 %. One step = one
batch through ratio_match(0.8, mutual) -> keypoint hand-off ->
ransac_fundamental(RansacConfig::default), as one plan (two captured
hipGraphs) over device-resident inputs.

The timed path ranks distances with BF16 MFMA (FP8 e4m3 with --precision fp8)
and re-scores every decision its error bound cannot settle in FP32, so its
match lists and inlier masks equal the exact path's; the plain FP32 GPU path
is timed the same way on the same batches, and the results are compared.

Steps are issued as a stream of batches: four distinct seeded batches, each
with its own plan on its own HIP stream, consecutive steps rotating over
them, so one batch's matching and RANSAC run beside the previous batch's
RANSAC tail. Every step is a complete run over a whole batch. The line also
reports one stream issuing steps back to back, the latency of an isolated
step, the time with host-to-device input copies and result copies included,
the FP64 error of the low-precision ranking, and secondary lines for the full
scene (3663 pairs) and the 8192-feature scene (FP8). The CPU oracle (the
restated reference algorithm) is timed on a bounded sample on the host cores.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "skylens-stream-rs_amd"))

METRIC = "matched image pairs/sec (ratio_match+RANSAC), BF16/FP8 vs FP32 GPU, exact matches"
PEAK = {"bf16": 2500.0, "fp8": 5000.0}  # dense MFMA TFLOP/s, MI355X
FP64_PEAK = 78.6  # TFLOP/s, MI355X
HBM_PEAK = 8000.0  # GB/s
FP32_VALU_PEAK_TOPS = 157.3  # TFLOP/s (FMA = 2); the exact kernel issues 3 separate ops per term
# FP64 flop of one 8-point hypothesis (normalised A^T A of 8 points, Householder
# tridiagonalisation, Sturm count, inverse iteration, rank-2 projection) and of
# one Sampson error evaluation (matching.rs:507-573).
FLOP_8PT = 3000.0
FLOP_SAMPSON = 40.0

WORKLOADS = {
    "uneven": ("uneven batch: 264 views (3 cameras x 88 positions), features/view log-uniform "
               "[200,8000], first 4000 candidate pairs, pair inlier ratio 10-90%, ratio 0.8 "
               "mutual, ransac_fundamental default config", 4),
    "full": ("full scene: 240 views (3 cameras x 80 positions), features/view U[1500,4096], "
             "60% noisy copies of a 20k pool, 3663 candidate pairs, pair outlier ratio 20-60%", 1),
    "fixed": ("fixed scene: 240 views x 8192 features, 3663 candidate pairs", 3),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_scene(name, seed):
    import workloads

    make = {"uneven": workloads.uneven_scene, "full": workloads.full_scene,
            "fixed": workloads.fixed_scene}[name]
    return make(seed=seed)


class Batch:
    """A scene with its descriptors and keypoints resident in HBM."""

    def __init__(self, name, seed):
        import numpy as np
        import torch

        t = time.time()
        self.sc = build_scene(name, seed)
        self.d_desc = torch.from_numpy(self.sc.desc).cuda()
        self.d_kp = torch.from_numpy(np.ascontiguousarray(self.sc.kp)).cuda()
        self.P = len(self.sc.pairs)
        na = np.array([self.sc.offsets[i + 1] - self.sc.offsets[i] for i, _ in self.sc.pairs],
                      np.float64)
        nb = np.array([self.sc.offsets[j + 1] - self.sc.offsets[j] for _, j in self.sc.pairs],
                      np.float64)
        self.rows = float(na.sum())
        self.gemm_flop = float(np.sum(2.0 * 128.0 * na * nb))
        log(f"{name} seed {seed}: {len(self.sc.views)} views, {self.P} pairs, "
            f"{self.sc.offsets[-1]} features ({time.time() - t:.1f} s)")

    def plan(self, ctx, prec, run_ransac=True):
        import skylens_hip

        return skylens_hip.Plan(ctx, self.sc.offsets, self.d_desc.data_ptr(), self.d_kp.data_ptr(),
                                self.sc.pairs, 0.8, True, prec, run_ransac, 0.0)


def sync_all(plans):
    import torch

    for p in plans:
        p.sync()
    torch.cuda.synchronize()


def rotation_seconds(plans, steps, warmup):
    """Seconds for `steps` steps issued in rotation over the plans (each on its own
    stream), bracketed by synchronisation; `warmup` untimed runs per plan first."""
    for _ in range(warmup):
        for p in plans:
            p.run()
    sync_all(plans)
    t = time.perf_counter()
    for k in range(steps):
        plans[k % len(plans)].run()
    sync_all(plans)
    return time.perf_counter() - t


def step_detail(plan):
    """(isolated step ms (median of 3), matching ms, RANSAC ms, distance kernel ms)."""
    import numpy as np

    iso = []
    for _ in range(3):
        t = time.perf_counter()
        plan.run()
        plan.sync()
        iso.append((time.perf_counter() - t) * 1e3)
    plan.set_timing(True)
    plan.run()
    plan.sync()
    match_ms, ransac_ms = plan.stage_ms(1), plan.stage_ms(2)
    plan.set_timing(False)
    return float(np.median(iso)), match_ms, ransac_ms, plan.time_distance(3)


def same_results(a, b, P):
    """Match lists, Some/None, and for the Some pairs F bits and inlier masks are equal."""
    import numpy as np

    return (np.array_equal(a.match_len, b.match_len)
            and all(np.array_equal(a.matches(p), b.matches(p)) for p in range(P))
            and np.array_equal(a.found, b.found)
            and all(np.array_equal(a.F[p].view(np.uint64), b.F[p].view(np.uint64))
                    and np.array_equal(a.inlier_mask(p), b.inlier_mask(p))
                    for p in range(P) if a.found[p] == 1))


def fp64_error(batch, plan, sample=12):
    """max |2 key - d^2_fp64| over the top-2 keys of the largest and some random
    pairs, the pair error bounds, and the share of rows re-scored exactly."""
    import numpy as np

    from fp64_check import scene_errors

    st = plan.lowp_state()
    size = np.array([(batch.sc.offsets[i + 1] - batch.sc.offsets[i]) *
                     (batch.sc.offsets[j + 1] - batch.sc.offsets[j]) for i, j in batch.sc.pairs])
    rng = np.random.default_rng(11)
    pick = np.unique(np.concatenate([np.argsort(size)[-sample // 2:],
                                     rng.choice(batch.P, sample // 2, replace=False)]))
    rep = scene_errors(batch.sc, st, pick)
    rep["rescored_row_frac"] = round(st["queued"] / batch.rows, 5)
    return {k: (round(v, 8) if isinstance(v, float) else v) for k, v in rep.items()}


def ransac_roofline(batch, plan, ransac_ms):
    """FP64 work of the RANSAC stage from the hypotheses each pair drew:
    hypotheses x (8-point solve + n Sampson errors), against the FP64 peak, and
    the bytes those hypotheses' scoring passes read (two f64 points per
    correspondence), against HBM. Local optimisation and the LM refits are not
    counted, so this is a lower bound on the work done."""
    import numpy as np

    iters, _, _ = plan.ransac_profile()
    res = plan.results()
    n = res.match_len.astype(np.float64)
    it = iters.astype(np.float64)
    flop = float(np.sum(it * (FLOP_8PT + FLOP_SAMPSON * n)))
    byts = float(np.sum(it * n * 32.0))
    t = ransac_ms * 1e-3
    return {"bound": "fp64", "kernel": "ransac_kernel (RANSAC stage: hand-off, cost probe, "
            "ordering, ransac_kernel)", "achieved": round(flop / t / 1e12, 3), "peak": FP64_PEAK,
            "unit": "TFLOP/s", "frac": round(flop / t / 1e12 / FP64_PEAK, 5),
            "flop_per_launch": flop, "hypotheses": int(np.sum(it)),
            "scoring_bytes_per_launch": byts,
            "scoring_GBps": round(byts / t / 1e9, 1), "hbm_frac": round(byts / t / 1e9 / HBM_PEAK, 5),
            "stage_ms": round(ransac_ms, 3)}


def tile_comparison(batch, ctx, prec):
    """Distance-kernel ms of one launch over the batch with each fixed tile and
    with the context's tuned table (pairs grouped by tile, the groups' launches
    overlapped)."""
    import skylens_hip

    out = {}
    for t, name in enumerate(skylens_hip.TILE_NAMES):
        ctx.set_tile(t)
        p = batch.plan(ctx, prec, run_ransac=False)
        p.run()
        p.sync()
        out[name] = round(p.time_distance(3), 4)
        p.close()
    ctx.set_tile(-1)
    p = batch.plan(ctx, prec, run_ransac=False)
    p.run()
    p.sync()
    out["tuned"] = round(p.time_distance(3), 4)
    out["tuned_tiles"] = p.tiles()
    p.close()
    best = min(out[n] for n in skylens_hip.TILE_NAMES)
    out["tuned_vs_best_fixed"] = round(best / out["tuned"], 4)
    return out


def e2e_with_h2d(batch, plan, steps=3):
    """Pairs/s when each step copies the inputs from pinned host memory, runs the
    plan and copies every result back (the PCIe-inclusive rate)."""
    import torch

    pin_d = torch.from_numpy(batch.sc.desc).pin_memory()
    pin_k = torch.from_numpy(batch.sc.kp.copy()).pin_memory()
    times = []
    for _ in range(steps + 1):
        torch.cuda.synchronize()
        t = time.perf_counter()
        batch.d_desc.copy_(pin_d, non_blocking=True)
        batch.d_kp.copy_(pin_k, non_blocking=True)
        torch.cuda.synchronize()
        plan.run()
        plan.results()
        times.append(time.perf_counter() - t)
    ms = 1e3 * min(times[1:])
    return {"ms_per_step": round(ms, 3), "pairs_per_s": round(batch.P / ms * 1e3, 2),
            "input_bytes": int(batch.sc.desc.nbytes + batch.sc.kp.nbytes)}


def secondary(name, precs, ctx):
    """One batch of `name`: isolated single-stream step per precision, identical
    results to FP32, FP64 error of the low-precision ranking, distance-kernel
    roofline."""
    b = Batch(name, WORKLOADS[name][1])
    out = {"workload": WORKLOADS[name][0], "pairs": b.P}
    ref = None
    res = {}
    for prec in precs + ["fp32"]:
        plan = b.plan(ctx, prec)
        plan.run()
        iso, m_ms, r_ms, k_ms = step_detail(plan)
        r = plan.results()
        d = {"isolated_step_ms": round(iso, 3), "pairs_per_s": round(b.P / iso * 1e3, 2),
             "match_ms": round(m_ms, 3), "ransac_ms": round(r_ms, 3),
             "distance_kernel_ms": round(k_ms, 4)}
        if prec in PEAK:
            d["tiles"] = plan.tiles()
            ach = b.gemm_flop / (k_ms * 1e-3) / 1e12
            d["distance_roofline"] = {"achieved": round(ach, 2), "peak": PEAK[prec],
                                      "unit": "TFLOP/s", "frac": round(ach / PEAK[prec], 4)}
            d["fp64_error"] = fp64_error(b, plan, sample=8 if name == "fixed" else 12)
        if prec == "fp32":
            ref = r
        res[prec] = (r, plan)
        out[prec] = d
        log(f"{name} {prec}: {d}")
    for prec in precs:
        out[prec]["identical_to_fp32_gpu"] = bool(same_results(res[prec][0], ref, b.P))
        out[prec]["speedup_vs_fp32_gpu"] = round(out["fp32"]["isolated_step_ms"] /
                                                 out[prec]["isolated_step_ms"], 3)
    ver = verified_input(b, res[precs[0]][1]) if name == "full" else None
    for _, p in res.values():
        p.close()
    out["verified_pairs"] = int(ref.found.sum())
    return out, b, {k: v[0] for k, v in res.items()}, ver


def verified_input(b, plan):
    """build_tracks' input from a plan's last run: every pair's verified matches
    (skh_plan_verified) and each view's keypoints."""
    import numpy as np

    vl, vij = plan.verified()
    base = plan.match_base
    pairs = [(int(i), int(j), vij[int(base[p]):int(base[p]) + int(vl[p])].astype(np.int64))
             for p, (i, j) in enumerate(b.sc.pairs) if vl[p]]
    kps = [b.sc.view_kp(v).astype(np.float64) for v in range(len(b.sc.offsets) - 1)]
    return pairs, kps


def tracks_line(ctx, ver, reps=3):
    """build_tracks (tracks.rs:190-508) on the full scene's verified matches (the
    RANSAC output of its BF16 plan), both policies: the call's wall time with
    the inputs and outputs in host memory (the device displacement filter
    included), tracks and stats."""
    import skylens_hip

    pairs, kps = ver
    packed = skylens_hip.TrackInput(pairs, kps)
    out = {"workload": f"full scene verified matches: {len(pairs)} pairs, "
                       f"{int(packed.m_off[-1])} matches, {len(kps)} views, {packed.n} keypoints"}
    for name, pol in (("split", skylens_hip.TRACKS_SPLIT), ("drop", skylens_hip.TRACKS_DROP)):
        cfg = skylens_hip.TrackConfig.default(policy=pol)
        ctx.build_tracks_packed(packed, cfg)
        t = time.perf_counter()
        for _ in range(reps):
            nt, st = ctx.build_tracks_packed(packed, cfg)
        out[name] = {"ms": round((time.perf_counter() - t) / reps * 1e3, 3), **st}
    return out, (pairs, kps)


def tracks_cpu(ver, gpu):
    """The oracle's build_tracks (one host thread) on the same input: time and
    whether its tracks and stats equal the device path's."""
    import skylens_hip
    import oracle_bind as ob

    pairs, kps = ver
    t = time.perf_counter()
    o_tracks, o_st = ob.build_tracks(pairs, kps, ob.track_cfg("split"))
    ms = (time.perf_counter() - t) * 1e3
    g_tracks, g_st = skylens_hip.default_context().build_tracks(pairs, kps)
    return {"value": round(ms, 1), "unit": "ms", "cores": 1, "kind": "port",
            "sample": "the full scene's verified matches through the track oracle (Split)",
            "identical_to_gpu": bool(o_tracks == g_tracks and o_st == g_st),
            "gpu_split_ms": gpu["split"]["ms"]}


def single_pair_latency(ctx):
    """The reference's two timing cases on the GPU through the single-pair C-ABI
    (host buffers in, results out: PCIe included): ratio_match on 7300 x 7300
    descriptors and ransac_fundamental on 4000 correspondences (50 % inliers)."""
    import numpy as np

    import skylens_hip
    import workloads

    a, b = workloads.scene_pair_descriptors(seed=21, na=7300, nb=7300)
    x1, x2 = workloads.two_view_correspondences(4000, 0.5, 0.5, seed=4242)
    out = {"ratio_match_7300x7300_ms": {}}
    for prec in ("bf16", "fp8", "fp32"):
        ts = []
        for _ in range(4):
            t = time.perf_counter()
            m = ctx.ratio_match(a, b, 0.8, True, prec)
            ts.append((time.perf_counter() - t) * 1e3)
        out["ratio_match_7300x7300_ms"][prec] = round(float(np.median(ts[1:])), 3)
    out["ratio_match_7300x7300_matches"] = len(m)
    ts = []
    for _ in range(4):
        t = time.perf_counter()
        fit = ctx.ransac_fundamental(x1, x2, skylens_hip.RansacConfig.default())
        ts.append((time.perf_counter() - t) * 1e3)
    out["ransac_4000_ms"] = round(float(np.median(ts[1:])), 3)
    out["ransac_4000_hypotheses"] = ctx.last_ransac_iters
    out["ransac_4000_inliers"] = int(fit.inliers.sum()) if fit is not None else 0
    return out, (a, b, x1, x2, fit)


def detector_line(n_ctx=4, reps=12):
    """detect_and_describe (features.rs:594-648) on 1920 x 1080 textured images
    resident in HBM (the reference's timing case, features.rs:1196-1208, is
    one 1920 x 1080 image in <= 0.4 s on the CPU): one context issuing images
    back to back, and n_ctx contexts on their own streams and host threads,
    one image each in flight."""
    import threading

    import numpy as np
    import torch

    import skylens_hip
    import workloads

    W, H = 1920, 1080
    cfg = skylens_hip.DetectorConfig.default()
    cap = int(cfg.max_features)
    imgs = [workloads.textured_image(W, H, 31 + k) for k in range(n_ctx)]
    d_imgs = [torch.from_numpy(i).cuda() for i in imgs]
    ctxs = [skylens_hip.Context(0) for _ in range(n_ctx)]
    d_kp = [torch.zeros((cap, 5), dtype=torch.float32, device="cuda") for _ in range(n_ctx)]
    d_desc = [torch.zeros((cap, 128), dtype=torch.float32, device="cuda") for _ in range(n_ctx)]
    torch.cuda.synchronize()

    def run(k, times):
        n = 0
        for _ in range(times):
            n = ctxs[k].detect_and_describe_dev(d_imgs[k].data_ptr(), W, H, cfg,
                                                d_kp[k].data_ptr(), d_desc[k].data_ptr(), cap)
        torch.cuda.synchronize()
        return n

    counts = [run(k, 2) for k in range(n_ctx)]  # warm-up, buffers grown
    t = time.perf_counter()
    run(0, reps)
    one = (time.perf_counter() - t) / reps
    ths = [threading.Thread(target=run, args=(k, reps)) for k in range(n_ctx)]
    t = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    multi = time.perf_counter() - t
    t = time.perf_counter()
    hk, hd = ctxs[0].detect_and_describe(imgs[0], cfg)  # host image in, results out
    host_ms = (time.perf_counter() - t) * 1e3
    out = {"workload": f"detect_and_describe, {W} x {H} textured images, default DetectorConfig",
           "features_per_image": counts, "ms_per_image_one_stream": round(one * 1e3, 3),
           "images_per_s_one_stream": round(1.0 / one, 2),
           "images_per_s": round(n_ctx * reps / multi, 2), "streams": n_ctx,
           "ms_per_image_host_buffers": round(host_ms, 3)}
    for c in ctxs:
        c.close()
    return out, (imgs[0], hk, hd)


def detector_cpu(det):
    """The detector oracle on the first image, single-threaded, and its
    bit-for-bit agreement with the GPU's features."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob

    img, hk, hd = det
    t = time.perf_counter()
    ok, od = ob.detect_and_describe(img)
    sec = time.perf_counter() - t
    same = (ok.shape == hk.shape and np.array_equal(ok.view(np.uint32), hk.view(np.uint32))
            and np.array_equal(od.view(np.uint32), hd.view(np.uint32)))
    return {"value": round(1.0 / sec, 3), "unit": "images/s", "cores": 1, "kind": "port",
            "sample": "one 1920 x 1080 image through the detector oracle on one host thread",
            "ms_per_image": round(sec * 1e3, 1), "features": int(ok.shape[0]),
            "identical_to_gpu": bool(same)}


def cpu_baseline(batch, gpu_res, threads, seconds, single, fixed):
    """The oracle (C restatement of the reference's CPU path) on a sample of the
    headline batch's pairs: ratio_match then ransac_fundamental per pair, pairs
    spread over `threads` host threads (the reference bench runs RANSAC par_iter
    over pairs; ctypes releases the GIL). Also: the reference's two single-pair
    timing cases on one thread, and the Frobenius distance between the oracle's
    and the GPU's F on the sampled pairs (FP64 refit on the CPU)."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob

    def one(sc, p):
        i, j = (int(v) for v in sc.pairs[p])
        m = ob.ratio_match(sc.view_desc(i), sc.view_desc(j), 0.8, True)
        mm = np.array(m, dtype=np.int64).reshape(-1, 2)
        x1 = sc.view_kp(i)[mm[:, 0]].astype(np.float64)
        x2 = sc.view_kp(j)[mm[:, 1]].astype(np.float64)
        return p, ob.ransac_fundamental(x1, x2)

    def frob(sc, res, picks):
        with ThreadPoolExecutor(threads) as ex:
            outs = list(ex.map(lambda p: one(sc, p), picks))
        err, masks = 0.0, True
        for p, r in outs:
            if (r is not None) != (res.found[p] == 1):
                return None, False
            if r is not None:
                err = max(err, float(np.linalg.norm(r[0] - res.F[p])))
                masks &= bool(np.array_equal(r[1], res.inlier_mask(p)))
        return err, masks

    sc, P = batch.sc, batch.P
    rng = np.random.default_rng(0)
    cal = rng.choice(P, size=min(P, 2 * threads), replace=False)
    t = time.time()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda p: one(sc, p), cal))
    per_pair = (time.time() - t) / len(cal)
    k = int(max(len(cal), min(P, seconds / max(per_pair, 1e-6))))
    sample = rng.choice(P, size=k, replace=False)
    t = time.time()
    with ThreadPoolExecutor(threads) as ex:
        outs = list(ex.map(lambda p: one(sc, p), sample))
    t_cpu = time.time() - t
    err, masks = 0.0, True
    for p, r in outs:
        if r is not None and gpu_res.found[p] == 1:
            err = max(err, float(np.linalg.norm(r[0] - gpu_res.F[p])))
            masks &= bool(np.array_equal(r[1], gpu_res.inlier_mask(p)))
    a, b, x1, x2, fit = single
    t = time.perf_counter()
    ob.ratio_match(a, b, 0.8, True)
    rm_ms = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    r4 = ob.ransac_fundamental(x1, x2)
    rs_ms = (time.perf_counter() - t) * 1e3
    out = {"value": round(k / t_cpu, 3), "unit": "pairs/s", "cores": threads, "kind": "port",
           "sample": f"{k} of the {P} pairs of the first headline batch, pairs spread over "
                     f"{threads} host threads, each pair's ratio_match and ransac_fundamental on "
                     f"one thread ({t_cpu:.1f} s); the reference bench instead runs ratio_match "
                     "pair by pair with rayon inside each pair",
           "f_frobenius_vs_gpu_max": err, "inlier_masks_equal": masks,
           "single_thread_ratio_match_7300x7300_ms": round(rm_ms, 1),
           "single_thread_ransac_4000_ms": round(rs_ms, 1),
           "ransac_4000_f_frobenius_vs_gpu": (float(np.linalg.norm(r4[0] - fit.f))
                                              if r4 is not None and fit is not None else None)}
    if fixed is not None:
        fb, fres = fixed
        picks = np.random.default_rng(1).choice(fb.P, size=min(fb.P, threads), replace=False)
        e, m = frob(fb.sc, fres, picks)
        out["fixed_scene_fp8"] = {"pairs_checked": len(picks), "f_frobenius_vs_gpu_max": e,
                                  "inlier_masks_equal": m}
    log(f"cpu oracle: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="uneven", choices=list(WORKLOADS))
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp8", "fp32"])
    ap.add_argument("--streams", type=int, default=4, choices=[1, 2, 3, 4],
                    help="distinct batches, each with its plan on its own stream, that "
                         "consecutive steps rotate over; 1 for profiler runs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the tile autotuner (every pair on the fixed default tile)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the full-scene and 8192-feature lines and the single-pair latencies")
    ap.add_argument("--kernel-only", action="store_true",
                    help="build the first batch's plan, run it once, then launch the distance "
                         "kernel 3 times (for rocprofv3 counter runs); prints no JSON")
    ap.add_argument("--detector-only", action="store_true",
                    help="run only the detector line (1920 x 1080 images) and print it; for "
                         "profiler runs")
    ap.add_argument("--tracks-only", action="store_true",
                    help="run the full scene's BF16 plan once and the build_tracks line on its "
                         "verified matches, print it; for profiler runs")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="run a plain device-to-device copy of BYTES three times (the reference "
                         "kernel for HBM counter calibration); prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path is measured on "
              "one MI355X and its multi-GPU sharding is a design note, not a run.", flush=True)
        sys.exit(2)

    import numpy as np
    import torch

    import skylens_hip

    if args.copy_calibration:
        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {4 * n} bytes 3 times")
        return

    if args.detector_only:
        out, _ = detector_line(n_ctx=args.streams)
        print(json.dumps({"detector": out}), flush=True)
        return

    if args.tracks_only:
        ctx = skylens_hip.Context(0)
        b = Batch("full", WORKLOADS["full"][1])
        plan = b.plan(ctx, "bf16")
        plan.run()
        ver = verified_input(b, plan)
        plan.close()
        out, _ = tracks_line(ctx, ver)
        print(json.dumps({"tracks": out}), flush=True)
        return

    S = args.streams
    desc, seed0 = WORKLOADS[args.workload]
    ctxs = [skylens_hip.Context(0) for _ in range(S)]
    tables = {}
    if not args.no_tune:
        # the tile autotuner: fastest tile per size class, loaded into every context
        for p in ("bf16", "fp8"):
            if args.precision == p or not args.no_secondary:
                t = time.time()
                tables[p], _ = ctxs[0].autotune(p)
                for c in ctxs[1:]:
                    c.set_tile_table(p, tables[p])
                log(f"autotune {p}: {tables[p].tolist()} ({time.time() - t:.1f} s)")
    batches = [Batch(args.workload, seed0 + s) for s in range(S)]
    torch.cuda.synchronize()
    prec = args.precision
    plans = [b.plan(c, prec) for b, c in zip(batches, ctxs)]
    log(f"{S} plans {prec} built")
    if args.kernel_only:
        plans[0].run()
        plans[0].sync()
        log(f"distance kernel {plans[0].time_distance(3):.3f} ms")
        for p in plans:
            p.close()
        return

    P = batches[0].P
    el = rotation_seconds(plans, args.steps, args.warmup)
    value = P * args.steps / el
    ms_step = el / args.steps * 1e3
    res = [p.results() for p in plans]
    el1 = rotation_seconds(plans[:1], args.steps, 1)
    iso, match_ms, ransac_ms, kern_ms = step_detail(plans[0])
    n_found = int(np.sum(res[0].found))
    n_matches = int(np.sum(res[0].match_len))
    log(f"{prec}: {ms_step:.2f} ms/step over {S} streams ({value:.0f} pairs/s); one stream "
        f"{el1 / args.steps * 1e3:.2f} ms/step; isolated {iso:.2f} ms ({match_ms:.2f} match, "
        f"{ransac_ms:.2f} ransac, distance kernel {kern_ms:.3f}); {n_matches} matches, "
        f"{n_found}/{P} verified")
    extra = {"batches": S,
             "steps_issue": (f"consecutive steps rotate over {S} distinct batches, each with its "
                             f"plan on its own stream" if S > 1 else "one plan, one stream"),
             "single_stream_pairs_per_s": round(P * args.steps / el1, 2),
             "single_stream_ms_per_step": round(el1 / args.steps * 1e3, 3),
             "isolated_step_ms": round(iso, 3)}
    if prec != "fp32":
        extra["fp64_error"] = fp64_error(batches[0], plans[0])
        log(f"fp64 error: {extra['fp64_error']}")
        if prec in tables:
            extra["tile_table"] = [skylens_hip.TILE_NAMES[k] for k in tables[prec]]
            extra["tiles"] = tile_comparison(batches[0], ctxs[0], prec)
            log(f"tiles: {extra['tiles']}")
    rroof = ransac_roofline(batches[0], plans[0], ransac_ms)
    log(f"ransac roofline: {rroof}")
    extra["e2e_with_h2d"] = e2e_with_h2d(batches[0], plans[0])
    log(f"with H2D: {extra['e2e_with_h2d']}")

    if prec != "fp32":
        plans32 = [b.plan(c, "fp32") for b, c in zip(batches, ctxs)]
        el32 = rotation_seconds(plans32, S, 1)
        res32 = [p.results() for p in plans32]
        same = all(same_results(r, r32, P) for r, r32 in zip(res, res32))
        v32 = P * S / el32
        iso32, m32, r32, k32 = step_detail(plans32[0])
        extra.update({"fp32_gpu_pairs_per_s": round(v32, 2),
                      "speedup_vs_fp32_gpu": round(value / v32, 3),
                      "fp32_ms_per_step": round(el32 / S * 1e3, 2),
                      "fp32_isolated_step_ms": round(iso32, 3),
                      "isolated_speedup_vs_fp32_gpu": round(iso32 / iso, 3),
                      "fp32_distance_kernel_ms": round(k32, 3),
                      "identical_to_fp32_gpu": bool(same)})
        log(f"fp32: {el32 / S * 1e3:.2f} ms/step ({v32:.0f} pairs/s), isolated {iso32:.2f} ms, "
            f"identical={same}")
        for p in plans32:
            p.close()

    gemm = batches[0].gemm_flop
    achieved = gemm / (kern_ms * 1e-3) / 1e12
    peak = PEAK.get(prec, FP32_VALU_PEAK_TOPS)
    roofline = {"bound": "mfma" if prec in PEAK else "valu",
                "kernel": "mfma_rank_kernel" if prec in PEAK else "exact_state_kernel",
                "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": None,
                "flop_per_launch": gemm, "launch_ms": round(kern_ms, 4)}
    prof = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{prec}.json")
    if os.path.exists(prof):
        with open(prof) as f:
            roofline["traffic"] = json.load(f).get("hbm_bytes_per_launch")
    for p in plans[1:]:
        p.close()
    for b in batches[1:]:
        del b.d_desc, b.d_kp

    fixed = None
    single = None
    det = None
    tracks_in = None
    if not args.no_secondary:
        sec = {}
        sec["full_scene"], fb, _, ver = secondary("full", ["bf16", "fp8"], ctxs[0])
        del fb
        sec["tracks"], tracks_in = tracks_line(ctxs[0], ver)
        log(f"tracks: {sec['tracks']}")
        sec["fixed_scene_fp8"], xb, xres, _ = secondary("fixed", ["fp8"], ctxs[0])
        fixed = (xb, xres["fp8"])
        sec["single_pair_latency"], single = single_pair_latency(ctxs[0])
        log(f"single pair: {sec['single_pair_latency']}")
        sec["detector"], det = detector_line()
        log(f"detector: {sec['detector']}")
        extra["secondary"] = sec

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        if single is None:
            _, single = single_pair_latency(ctxs[0])
        cpu = cpu_baseline(batches[0], res[0], threads, args.cpu_seconds, single, fixed)
        if det is not None:
            cpu["detector"] = detector_cpu(det)
            log(f"detector cpu: {cpu['detector']}")
        if tracks_in is not None:
            cpu["tracks"] = tracks_cpu(tracks_in, sec["tracks"])
            log(f"tracks cpu: {cpu['tracks']}")

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": prec, "data": "synthetic",
        "config": {"workload": desc, "pairs": P, "views": len(batches[0].sc.views),
                   "features": int(batches[0].sc.offsets[-1]), "precision": prec,
                   "matches": n_matches, "verified_pairs": n_found,
                   "match_ms": round(match_ms, 3), "ransac_ms": round(ransac_ms, 3), **extra},
        "roofline": roofline,
        "roofline_ransac": rroof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    plans[0].close()
    for c in ctxs:
        c.close()


if __name__ == "__main__":
    main()
