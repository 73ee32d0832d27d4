"""Benchmark: plane-mode PatchMatch MVS on one MI355X.

Default workload (`--workload batch`): the batched mixed-precision run -- 48
reference views of uneven resolution (640x480 .. 1920x1440, 66 Mpx in total)
from a 64-camera orbit around a sphere and two boxes (synthetic, analytic
depth), each against 4-10 sources, 11x11 window with stride 2, top-4 views,
5 red-black iterations, images stored as FP16 with FP32 arithmetic, all
references in one persistent HIP-graph launch per phase.  One step = one
replay of that graph.  Inputs are resident in HBM before the timed region.

metric: Gpix.view.iter/s = sum over references of P x V x iterations / time /
1e9 (whole job), and the median depth error against the analytic truth.

Also on the line: the batch kernel's roofline (algorithmic FLOPs per launch /
mean launch time from HIP events on the library's stream; HBM bytes per
launch from the committed rocprofv3 profile it names), the reference CPU path
(compiled from the reference's own sources) on a bounded crop of one of the
workload's references with the GPU's error on the same crop, and as extra
keys the plain FP32 GPU path (one launch per phase and view, FP32 images) on
the same 48 jobs, the 9-view 1920x1440 depth-map workload, and reference mode
(the bit-exact drop-in path: time per 1280x960 view, bit-exactness against
the reference's own code, FP64 roofline).

Other workloads: depthmaps (the 9-view job as the headline), adjoint,
reference, copy (counter calibration).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "PatchMatch Gpix*view*iter/s + median depth error vs analytic truth, GPU vs CPU"
PEAK_FP32_TFLOPS = 157.3    # MI355X FP32 vector peak
PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 vector peak (reference mode)
PEAK_HBM_GBS = 8000.0
PEAK_LDS_GBS = 78643.2      # LDS: 128 B per clock per CU x 256 CUs x 2.4 GHz
FLOP_PER_SAMPLE = 20        # warp 6 + bilinear 9 + accumulate 5 per window sample
HYPS_PER_ITER = 8           # 4 propagated + 3 perturbed + 1 random
SAMPLES_11x11_S2 = 36
# reference mode, per sample pair: source bilinear 12 (u, fx, 1 - fx, three
# lerps of 3) + the sums of the source values, their squares and the cross
# products 5; the reference patch's own sums are per pixel
REF_FLOP_PER_PAIR = 17


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="batch",
                    choices=["batch", "depthmaps", "adjoint", "reference", "copy"],
                    help="batch: 48 references of uneven size in one captured batch, FP16 images "
                         "(default); depthmaps: the 9-view 1920x1440 job; adjoint: cost + gradients "
                         "of a 960x720 depth map against 6 misaligned sources, and the 10-step "
                         "polish; reference: reference-mode PatchMatchStereo + RefineSubpixel of one "
                         "1280x960 view, the reference's CPU code on a 320x240 view as baseline and "
                         "bit-exactness check; copy: a plain 1 GiB device copy (calibrates rocprofv3 "
                         "byte counters)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1440)
    ap.add_argument("--iterations", type=int, default=5)
    ap.add_argument("--storage", default=None, choices=["f32", "f16", "bf16", "u8"],
                    help="image storage (batch: f16; depthmaps: u8 for --images u8, f32 for "
                         "--images float)")
    ap.add_argument("--images", default="u8", choices=["float", "u8"],
                    help="depthmaps: the renderer's images quantized to 8 bits (what a camera and "
                         "the reference deliver), or its float images")
    ap.add_argument("--view-phased", type=int, default=-1,
                    help="depthmaps: leading iterations run view-phased; -1: automatic")
    ap.add_argument("--tile-schedule", type=int, default=-1,
                    help="workgroup tile schedule; -1 = autotuned with the per-phase occupancy "
                         "(untimed set-up)")
    ap.add_argument("--check-size", default="320x240",
                    help="reference mode: size (WxH) of the view run by the reference's CPU code "
                         "as the baseline and compared with the GPU bit for bit")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target duration of the CPU-baseline sample (0 = skip)")
    ap.add_argument("--extras", type=int, default=1,
                    help="batch: also run the plain FP32 path, the depth-map workload and "
                         "reference mode as extra keys (0 = headline only)")
    return ap.parse_args()


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def traffic_entry(key, kind="traffic"):
    """A counter figure per launch for `key` from profiles/hbm_traffic.txt
    (`kind`: traffic = HBM bytes, lds = LDS bytes loaded, valu_issue = VALU
    issue fraction), with the profile it was measured in."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if os.path.exists(path):
        for line in open(path):
            f = line.split()
            if len(f) == 4 and f[0] == kind and f[1] == key:
                return float(f[2]), "profiles/" + f[3]
    return None, None


def roofline(flop_per_launch, mean_ms, kernel, traffic_key, peak=PEAK_FP32_TFLOPS, **extra):
    achieved = flop_per_launch / (mean_ms * 1e-3) / 1e12 if mean_ms else None
    traffic, src = traffic_entry(traffic_key)
    r = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
         "frac": achieved / peak if achieved else None, "traffic": traffic,
         "traffic_source": src, "kernel": kernel, "mean_launch_ms": mean_ms,
         "flop_per_launch": flop_per_launch}
    if traffic and mean_ms:
        r["hbm_gbs"] = traffic / (mean_ms * 1e-3) / 1e9
        r["hbm_frac"] = r["hbm_gbs"] / PEAK_HBM_GBS
    lds, _ = traffic_entry(traffic_key, "lds")
    if lds and mean_ms:
        r["lds_gbs"] = lds / (mean_ms * 1e-3) / 1e9
        r["lds_frac"] = r["lds_gbs"] / PEAK_LDS_GBS
    issue, _ = traffic_entry(traffic_key, "valu_issue")
    if issue:
        r["valu_issue"] = issue
    r.update(extra)
    return r


def cpu_baseline(sc, ref, srcs, iterations, target_s, gpu_depth=None):
    """The reference's PatchMatchStereo (fronto-parallel, 11x11, all sources,
    `iterations` x 4 directions, every host core) on a centred crop of view
    `ref` (8-bit images, the reference's only format), sized to ~target_s.
    With gpu_depth (the GPU's map of `ref`), the GPU's median error on the
    same crop is reported beside the CPU's."""
    import oracle
    from recon_scanner_amd import _native as N
    if not oracle.reference_available():
        return None
    cores = len(os.sched_getaffinity(0))
    u8 = {i: sc.u8(i) for i in [ref] + list(srcs)}
    H, W = u8[ref].shape
    V = len(srcs)

    def run(cw, ch, iters):
        x0, y0 = (W - cw) // 2, (H - ch) // 2
        crop = np.ascontiguousarray(u8[ref][y0:y0 + ch, x0:x0 + cw])
        c0 = sc.cams[ref]
        rc = N.Camera.make(c0.fx, c0.fy, c0.cx - x0, c0.cy - y0, cw, ch, c0.R_np(), c0.t_np())
        cfg = N.config_default()
        cfg.patch_size, cfg.iterations, cfg.num_threads = 11, iters, cores
        cfg.min_depth_m, cfg.max_depth_m = sc.min_depth, sc.max_depth
        cfg.num_views = V
        poses = [rc] + [sc.cams[s] for s in srcs]
        t = time.perf_counter()
        f, d, s, c = oracle.ref_patchmatch(crop, [u8[s] for s in srcs], rc, poses, poses,
                                           list(range(1, V + 1)), cfg)
        return time.perf_counter() - t, d, (x0, y0)

    # calibrate on a crop tall enough for every host thread (the reference
    # splits rows into min(threads, rows / 16) blocks), then size the sample
    ch0 = min(H, max(120, 16 * cores))
    cw0 = min(W, max(160, ch0 * 4 // 3))
    dt, _, _ = run(cw0, ch0, 1)
    per_unit = dt / (cw0 * ch0 * V)
    area = max(64 * 48, min(W * H, int(target_s / (per_unit * V * iterations))))
    ch = max(48, min(H, int((area * 3 / 4) ** 0.5)))
    cw = max(64, min(W, int(ch * 4 / 3)))
    dt, d, (x0, y0) = run(cw, ch, iterations)
    truth = sc.depths[ref][y0:y0 + ch, x0:x0 + cw]
    m = np.isfinite(d) & (truth > 0)
    out = {"value": cw * ch * V * iterations / dt / 1e9, "unit": "Gpix*view*iter/s", "cores": cores,
           "kind": "reference",
           "sample": f"reference PatchMatchStereo compiled from its sources (fronto-parallel, 11x11, "
                     f"{V} views, {iterations} iterations x 4 directions, num_threads={cores}) on a "
                     f"centred {cw}x{ch} crop of reference view {ref} (8-bit images); {dt:.1f} s",
           "seconds": dt, "median_depth_error_mm": float(np.median(np.abs(d[m] - truth[m]))) * 1e3}
    if gpu_depth is not None:
        g = gpu_depth[y0:y0 + ch, x0:x0 + cw]
        mg = np.isfinite(g) & (truth > 0)
        both = mg & m
        out["gpu_median_depth_error_mm_same_crop"] = float(np.median(np.abs(g[mg] - truth[mg]))) * 1e3
        out["median_depth_error_mm_pixels_valid_on_both"] = {
            "cpu": float(np.median(np.abs(d[both] - truth[both]))) * 1e3,
            "gpu": float(np.median(np.abs(g[both] - truth[both]))) * 1e3,
            "pixels": int(both.sum())}
    return out


def base_record(args, dt, steps, workload, **config):
    return {"n_gpus": 1, "steps": steps, "warmup": args.warmup,
            "ms_per_step": dt / steps * 1000.0, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "config": dict(workload=workload, **config)}


def depth_errors(pm, sc, refs):
    """Median relative and absolute (mm) |Z - Z_true| over the valid pixels of
    the maps of `refs`, and the maps."""
    rel, ab, maps = [], [], {}
    for r in refs:
        d = pm.depth_map(r)["depth"]
        maps[r] = d
        m = np.isfinite(d) & (sc.depths[r] > 0)
        e = np.abs(d[m] - sc.depths[r][m])
        ab.append(e)
        rel.append(e / sc.depths[r][m])
    return float(np.median(np.concatenate(rel))), float(np.median(np.concatenate(ab))) * 1e3, maps


def batch_flop(sc, jobs, iterations):
    """Algorithmic FLOPs of one batched run: per reference P x V x S x 20 for
    the random init (1 hypothesis per pixel) and 8 hypotheses per pixel per
    iteration."""
    pv = sum(sc.cams[r].width * sc.cams[r].height * len(s) for r, s in jobs)
    return pv * SAMPLES_11x11_S2 * FLOP_PER_SAMPLE * (1 + HYPS_PER_ITER * iterations)


def run_batch(args):
    from recon_scanner_amd import plane, scenes
    storage = args.storage or "f16"
    t0 = time.time()
    sc, jobs = scenes.orbit_batch(num_refs=48, num_cams=64, seed=0)
    log(f"orbit rendered in {time.time() - t0:.1f} s")
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=args.iterations,
                             min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0,
                             tile_schedule=args.tile_schedule)
    refs = [r for r, _ in jobs]
    units = sum(sc.cams[r].width * sc.cams[r].height * len(s) for r, s in jobs) * args.iterations
    mpx = sum(sc.cams[r].width * sc.cams[r].height for r in refs) / 1e6
    # batched mixed-precision path (the headline)
    pm = plane.PlaneMVS(sc.cams, sc.images, storage=storage)
    b = pm.batch(jobs, p)
    for _ in range(args.warmup):
        b.run()
    pm.synchronize()
    log(f"batch built and warmed at {time.time() - t0:.1f} s")
    pm.timing(enable=True, reset=True)
    t = time.perf_counter()
    for _ in range(args.steps):
        b.run()
    pm.synchronize()
    dt = time.perf_counter() - t
    tim = pm.timing(enable=False, reset=True)
    rel_b, abs_b, maps_b = depth_errors(pm, sc, refs)
    schedule = b.schedule
    b.close()
    graph_ms = tim["batch"][0] / max(1, tim["batch"][1])
    launches = 1 + 2 * args.iterations
    flop = batch_flop(sc, jobs, args.iterations)
    value = units * args.steps / dt / 1e9
    log(f"{args.steps} batched runs in {dt:.3f} s ({graph_ms:.1f} ms per graph)")
    out = {"metric": METRIC, "value": value, "unit": "Gpix*view*iter/s"}
    out.update(base_record(args, dt, args.steps,
                           f"batched mixed-precision PatchMatch: 48 references of uneven size "
                           f"(640x480..1920x1440, {mpx:.1f} Mpx) from a 64-camera orbit, 4-10 sources "
                           f"each, 11x11 stride 2, top-4, {args.iterations} red-black iterations, "
                           f"{storage.upper()} images / FP32 arithmetic, one HIP-graph launch per run",
                           refs=len(refs), megapixels=mpx, storage=storage, schedule=schedule,
                           iterations=args.iterations))
    out.update({"dtype": "fp32", "data": "synthetic (64-camera orbit around a sphere and two boxes, "
                                         "analytic depth, float images)",
                "median_depth_error_mm": abs_b, "median_rel_depth_error": rel_b,
                "depth_maps_per_s": len(refs) * args.steps / dt,
                "hypothesis_view_evals_per_s":
                    units / args.iterations * (1 + HYPS_PER_ITER * args.iterations) * args.steps / dt,
                "roofline": roofline(flop / launches, graph_ms / launches,
                                     "k_plane_batch (one persistent launch per phase: init + "
                                     f"{2 * args.iterations} half-sweeps, averaged)",
                                     f"batch-orbit48-{storage}", graph_ms=graph_ms,
                                     launches_per_graph=launches)})
    cpu = None
    if args.cpu_seconds > 0:
        try:
            r0, s0 = jobs[0]
            cpu = cpu_baseline(sc, r0, s0, args.iterations, args.cpu_seconds, maps_b[r0])
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline failed: {e}")
    out["cpu_baseline"] = cpu
    if not args.extras:
        print(json.dumps(out))
        return
    # the plain FP32 GPU path on the same jobs: FP32 images, one launch per
    # phase and reference view
    xsteps = max(1, min(args.steps, 3))
    pf = plane.PlaneMVS(sc.cams, sc.images, storage="f32")

    def plain():
        for r, srcs in jobs:
            pf.patchmatch(r, srcs, p)
    plain()
    pf.synchronize()
    t = time.perf_counter()
    for _ in range(xsteps):
        plain()
    pf.synchronize()
    dtp = time.perf_counter() - t
    rel_p, abs_p, _ = depth_errors(pf, sc, refs)
    plain_rate = units * xsteps / dtp / 1e9
    log(f"plain FP32 path: {dtp / xsteps * 1e3:.1f} ms per pass")
    out["plain_fp32"] = {"value": plain_rate, "unit": "Gpix*view*iter/s",
                         "depth_maps_per_s": len(refs) * xsteps / dtp,
                         "ms_per_pass": dtp / xsteps * 1e3, "steps": xsteps,
                         "median_depth_error_mm": abs_p, "median_rel_depth_error": rel_p}
    out["batched_median_rel_error_over_plain"] = rel_b / rel_p if rel_p > 0 else None
    tgt = {"batched_over_plain": value / plain_rate,
           "reading": "(ii) both GPU paths at least 50x the CPU reference's depth-map throughput; "
                      "reading (i), batched at least 50x the plain FP32 GPU path, is reported too"}
    if cpu:
        tgt["batched_over_cpu"] = value / cpu["value"]
        tgt["plain_over_cpu"] = plain_rate / cpu["value"]
        tgt["met_reading_ii"] = bool(tgt["batched_over_cpu"] >= 50 and tgt["plain_over_cpu"] >= 50)
    tgt["met_reading_i"] = bool(tgt["batched_over_plain"] >= 50)
    tgt["accuracy_gate_met"] = bool(abs(rel_b - rel_p) <= 1e-3 * rel_p)
    out["target_50x"] = tgt
    del pf
    try:
        out["depthmaps_1920x1440"] = run_depthmaps(args, steps=xsteps, headline=False)
    except Exception as e:
        log(f"depthmaps leg failed: {e}")
    try:
        out["reference_mode"] = run_reference(args, steps=xsteps, headline=False)
    except Exception as e:
        log(f"reference leg failed: {e}")
    print(json.dumps(out))


def run_adjoint(args):
    """Cost and gradients (per-pixel plane, 6 source poses) of a 960x720 depth
    map whose sources are misaligned by 0.5 deg / 1 mm (seeded), then the
    10-step gradient polish of a fresh PatchMatch result."""
    from recon_scanner_amd import plane, scenes
    sc = scenes.sphere_planes(width=960, height=720, focal=1250.0, n_views=7, supersample=2)
    rng = np.random.default_rng(5)
    cams = list(sc.cams)
    for k in range(1, 7):
        ax, tr = rng.normal(size=3), rng.normal(size=3)
        d = np.concatenate([np.radians(0.5) * ax / np.linalg.norm(ax), 1e-3 * tr / np.linalg.norm(tr)])
        cams[k] = scenes.perturb_camera(sc.cams[k], d)
    srcs = list(range(1, 7))
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=args.iterations,
                             min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0,
                             tile_schedule=args.tile_schedule)
    pm = plane.PlaneMVS(cams, sc.images, storage=args.storage or "f32")
    pm.patchmatch(0, srcs, p)
    for _ in range(args.warmup):
        pm.cost_grad(0, srcs, p)
    pm.timing(enable=True, reset=True)
    t = time.perf_counter()
    for _ in range(args.steps):
        pm.cost_grad(0, srcs, p)
    dt = time.perf_counter() - t
    tim = pm.timing(enable=False, reset=True)
    adj_ms = tim["adjoint"][0] / max(1, tim["adjoint"][1])
    P, V = 960 * 720, 6
    pm.timing(enable=True, reset=True)
    pm.polish(0, srcs, p, steps=10)
    pm.synchronize()
    pol_ms = pm.timing(enable=False, reset=True)["adjoint"][0]
    out = {"metric": "adjoint Gpix*view/s (cost + plane and pose gradients)",
           "value": P * V / (adj_ms * 1e-3) / 1e9, "unit": "Gpix*view/s", "dtype": "fp32",
           "data": "synthetic (rendered sphere + 2 tilted planes, poses misaligned 0.5 deg / 1 mm)"}
    out.update(base_record(args, dt, args.steps, "adjoint of a 960x720 depth map, 6 sources, "
                                                 "11x11 stride 2, top-4 (host copies included in "
                                                 "ms_per_step)", width=960, height=720, sources=6))
    out.update({"adjoint_kernel_ms": adj_ms, "polish_10_steps_ms": pol_ms})
    print(json.dumps(out))


def run_reference(args, steps=None, headline=True):
    """The drop-in path: reference-mode PatchMatchStereo + RefineSubpixel (the
    reference's algorithm, bit-exact) of one view of the sphere + planes scene
    against its 8 sources.  Throughput on the GPU at 1280x960 (or --width x
    --height); the reference's own CPU code (same config and num_threads, whose
    row-block RNG streams the GPU reproduces) on a 320x240 view (or
    --check-size) as the baseline, where the GPU's outputs are also checked
    against it bit for bit (the CPU takes ~20 s there on 16 cores)."""
    import oracle
    from recon_scanner_amd import mvs, scenes
    steps = steps or args.steps
    cores = len(os.sched_getaffinity(0))

    def problem(w, h):
        sc = scenes.sphere_planes(width=w, height=h, focal=2500.0 * w / 1920.0, n_views=9, supersample=2)
        cfg = mvs.MvsConfig(patch_size=11, iterations=args.iterations, num_threads=cores,
                            min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, num_views=8)
        return sc, [sc.u8(i) for i in range(sc.num_views)], cfg

    srcs = list(range(1, 9))

    def gpu_once(sc, u8, cfg):
        h, w = u8[0].shape
        poses = [mvs.CameraPose.from_camera(c) for c in sc.cams]
        intrs = [mvs.CameraIntrinsics.from_camera(c) for c in sc.cams]
        dm = mvs.DepthMap()
        filled = mvs.PatchMatchStereo(u8[0], w, h, [u8[i] for i in srcs], [w] * 8, [h] * 8, poses[0],
                                      poses, intrs[0], intrs, srcs, cfg, None, dm)
        refined = mvs.RefineSubpixel(u8[0], w, h, [u8[i] for i in srcs], [w] * 8, [h] * 8, poses[0],
                                     poses, intrs[0], intrs, srcs, cfg, dm)
        return filled, refined, dm

    w, h = (args.width, args.height) if (headline and args.width != 1920) else (1280, 960)
    sc, u8, cfg = problem(w, h)
    L, ctx = mvs.lib(), mvs.context()

    def timers(enable, reset):
        ms, cnt = np.zeros(8), np.zeros(8, np.int64)
        mvs.check(L.recon_mvs_hip_timing(ctx, int(enable), int(reset), ms.ctypes.data,
                                         cnt.ctypes.data), "timing")
        return ms, cnt

    for _ in range(args.warmup if headline else 1):
        gpu_once(sc, u8, cfg)
    timers(True, True)
    t = time.perf_counter()
    for _ in range(steps):
        filled, refined, dm = gpu_once(sc, u8, cfg)
    dt = time.perf_counter() - t
    ms, cnt = timers(False, True)
    units = w * h * 8 * args.iterations
    rec = {"metric": "reference-mode PatchMatchStereo px*view*iter/s (bit-exact with the reference)",
           "value": units * steps / dt / 1e9, "unit": "Gpix*view*iter/s", "dtype": "fp64"}
    rec.update(base_record(args, dt, steps, f"reference-mode PatchMatchStereo + RefineSubpixel, "
                           f"{w}x{h}, 8 sources, 11x11, {args.iterations} iterations x 4 directions, "
                           f"num_threads={cores}", width=w, height=h, iterations=args.iterations))
    rec["ms_per_view"] = dt / steps * 1e3
    truth = sc.depths[0]
    d = dm.depth.reshape(h, w)
    m = np.isfinite(d) & (truth > 0)
    rec["median_depth_error_mm"] = float(np.median(np.abs(d[m] - truth[m]))) * 1000.0
    # roofline of one directional pass (k_pm_pass and its k_pm_restart, timed
    # together by HIP events on the library's stream): nominal hypotheses per
    # pixel and pass H = 4 neighbours + random_search_count + restart_ratio,
    # each over 8 views x patch^2 sample pairs of REF_FLOP_PER_PAIR FP64 operations
    pass_ms = ms[0] / cnt[0] if cnt[0] else None
    hyps = 4 + cfg.random_search_count + cfg.restart_ratio
    flop = w * h * hyps * 8 * cfg.patch_size ** 2 * REF_FLOP_PER_PAIR
    rec["roofline"] = roofline(flop, pass_ms, "one directional pass: k_pm_pass + k_pm_restart",
                               f"reference-{w}x{h}", peak=PEAK_FP64_TFLOPS,
                               hypotheses_per_pixel=hyps)
    if args.cpu_seconds > 0 and oracle.reference_available():
        cw, ch = (int(x) for x in args.check_size.lower().split("x"))
        sc, u8, cfg = problem(cw, ch)
        filled, refined, dm = gpu_once(sc, u8, cfg)
        c0 = time.perf_counter()
        rf, rd, rs, rc = oracle.ref_patchmatch(u8[0], [u8[i] for i in srcs], sc.cams[0], sc.cams,
                                               sc.cams, srcs, cfg)
        rn, rdr, rcr = oracle.ref_refine(u8[0], [u8[i] for i in srcs], sc.cams[0], sc.cams, sc.cams,
                                         srcs, cfg, rd, rs, rc)
        cdt = time.perf_counter() - c0
        exact = (rf == filled and rn == refined and
                 dm.depth.reshape(ch, cw).tobytes() == rdr.tobytes() and
                 dm.support_count.reshape(ch, cw).tobytes() == rs.tobytes() and
                 dm.confidence.reshape(ch, cw).tobytes() == rcr.tobytes())
        cpu_rate = cw * ch * 8 * args.iterations / cdt / 1e9
        rec["cpu_baseline"] = {"value": cpu_rate, "unit": "Gpix*view*iter/s", "cores": cores,
                               "kind": "reference",
                               "sample": f"reference PatchMatchStereo + RefineSubpixel compiled from "
                                         f"its sources, the same scene and config at {cw}x{ch}; {cdt:.1f} s",
                               "seconds": cdt}
        rec[f"bit_exact_with_reference_{cw}x{ch}"] = bool(exact)
        rec["speedup_vs_reference_cpu"] = rec["value"] / cpu_rate
    if headline:
        print(json.dumps(rec))
    return rec


def run_copy(args):
    """A plain device-to-device copy of 1 GiB, `steps` times: known bytes for
    calibrating FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    out = {"metric": "device copy GB/s", "value": 2 * 4 * n * args.steps / dt / 1e9, "unit": "GB/s",
           "dtype": "fp32", "data": "constant", "bytes_read_per_copy": 4 * n,
           "bytes_written_per_copy": 4 * n}
    out.update(base_record(args, dt, args.steps, "1 GiB device-to-device copy"))
    print(json.dumps(out))


def run_depthmaps(args, steps=None, headline=True):
    """The 9-view job: the sphere + two tilted planes scene at 1920x1440 (f =
    2500 px, 100 mm ring baseline); every view is a reference once with the
    other 8 as sources (random plane init + 5 red-black iterations, 11x11
    stride 2, top-4), then the geometric-consistency filter over the 9 maps.
    8-bit images in 8-bit storage by default (bit-identical to FP32 storage)."""
    from recon_scanner_amd import plane, scenes
    steps = steps or args.steps
    storage = args.storage if (headline and args.storage) else ("u8" if args.images == "u8" else "f32")
    t0 = time.time()
    W, H = (args.width, args.height) if headline else (1920, 1440)
    f = 2500.0 * W / 1920.0
    sc = scenes.sphere_planes(width=W, height=H, focal=f, n_views=9, supersample=2)
    log(f"scene rendered in {time.time() - t0:.1f} s")
    nv = sc.num_views
    views = list(range(nv))
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=args.iterations,
                             min_depth_m=sc.min_depth, max_depth_m=sc.max_depth,
                             value_scale=255.0 if args.images == "u8" else 1.0,
                             tile_schedule=args.tile_schedule,
                             view_phased_iterations=args.view_phased)
    imgs = sc.images if args.images == "float" else [sc.u8(i) for i in range(nv)]
    pm = plane.PlaneMVS(sc.cams, imgs, storage=storage)

    def step():
        for r in views:
            pm.patchmatch(r, [v for v in views if v != r], p)
        pm.geo_consistency(views, 0.01)

    if args.tile_schedule < 0:
        # untimed set-up: tile schedule and per-phase occupancy are tuned once
        # per (size, views, window, format), here on view 0
        pm.patchmatch(views[0], views[1:], p)
        pm.synchronize()
    for _ in range(args.warmup):
        step()
    pm.synchronize()
    pm.timing(enable=True, reset=True)
    t = time.perf_counter()
    for _ in range(steps):
        step()
    pm.synchronize()
    dt = time.perf_counter() - t
    tim = pm.timing(enable=False, reset=True)
    log(f"depthmaps: {steps} steps in {dt:.3f} s; kernels {tim}")
    P = W * H
    V = nv - 1
    units = nv * P * V * args.iterations
    hve = nv * P * V * (1 + HYPS_PER_ITER * args.iterations)
    sweep_ms, sweep_n = tim["sweep"]
    flop_launch = (P / 2) * HYPS_PER_ITER * V * SAMPLES_11x11_S2 * FLOP_PER_SAMPLE
    # accuracy: median |Z - Z_true| over consistency-kept pixels of all maps
    errs = []
    kept = 0
    for r in views:
        d = pm.depth_map(r, drop_inconsistent=2)["depth"]
        m = np.isfinite(d) & (sc.depths[r] > 0)
        kept += int(m.sum())
        errs.append(np.abs(d[m] - sc.depths[r][m]))
    med = float(np.median(np.concatenate(errs))) * 1000.0
    out = {"metric": METRIC, "value": units * steps / dt / 1e9, "unit": "Gpix*view*iter/s",
           "dtype": "fp32"}
    out.update(base_record(args, dt, steps, f"plane-mode PatchMatch, 9 views {W}x{H}, each view vs "
                           f"8 sources, 11x11 stride 2, top-4, {args.iterations} red-black iterations "
                           f"+ geometric-consistency filter over 9 depth maps",
                           views=nv, width=W, height=H, iterations=args.iterations, images=args.images,
                           storage=storage, view_phased_iterations=args.view_phased))
    out.update({"data": "synthetic (rendered sphere + 2 tilted planes, analytic depth"
                        + (", 8-bit images)" if args.images == "u8" else ", float images)"),
                "hypothesis_view_evals_per_s": hve * steps / dt, "median_depth_error_mm": med,
                "kept_pixels": kept,
                "roofline": roofline(flop_launch, sweep_ms / sweep_n if sweep_n else None,
                                     "half-sweep: k_plane_sweep (k_phase_prep + k_phase_view per "
                                     "source + k_phase_final in view-phased iterations)",
                                     f"depthmaps-{W}x{H}-{storage}")})
    if headline:
        cpu = None
        if args.cpu_seconds > 0:
            try:
                cpu = cpu_baseline(sc, 0, list(range(1, nv)), args.iterations, args.cpu_seconds,
                                   pm.depth_map(0)["depth"])
            except Exception as e:  # the baseline is reported, never required
                log(f"cpu baseline failed: {e}")
        out["cpu_baseline"] = cpu
        print(json.dumps(out))
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the sharding "
                                   "over GPUs is a design, not part of this build)"}))
        sys.exit(2)
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "adjoint":
        return run_adjoint(args)
    if args.workload == "reference":
        return run_reference(args)
    if args.workload == "copy":
        return run_copy(args)
    return run_depthmaps(args)


if __name__ == "__main__":
    main()
