"""Benchmark: plane-mode PatchMatch MVS on one MI355X.

Workload (one step): the 9-view sphere + two tilted planes scene at 1920x1440
(f = 2500 px, 100 mm ring baseline, synthetic, analytic depth).  Every view is
a reference once with the other 8 as sources: random plane init + 5 red-black
iterations, 11x11 window with stride 2 (36 samples), top-4 of 8 views, FP32
arithmetic; then the geometric-consistency filter across the 9 depth maps.
Inputs are resident in HBM before the timed region starts.

metric: Gpix.view.iter/s = 9 maps x P x V x iterations / time / 1e9 (whole job).
Also reported: hypothesis-view evaluations/s, median |Z - Z_true| of the
consistency-filtered maps, the sweep kernel's roofline (HIP events on the
library's stream) and the reference CPU path (compiled from the reference's
own sources) timed on a bounded crop of the same workload on this host.
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

PEAK_FP32_TFLOPS = 157.3    # MI355X FP32 vector peak
PEAK_HBM_GBS = 8000.0
FLOP_PER_SAMPLE = 20        # warp 6 + bilinear 9 + accumulate 5 per window sample
HYPS_PER_ITER = 8           # 4 propagated + 3 perturbed + 1 random
PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 vector peak (reference mode)
# reference mode, per sample pair: source bilinear 12 (u, fx, 1 - fx, three
# lerps of 3) + the sums of the source values, their squares and the cross
# products 5; the reference patch's own sums are per pixel
REF_FLOP_PER_PAIR = 17


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="depthmaps",
                    choices=["depthmaps", "batch", "adjoint", "reference", "copy"],
                    help="depthmaps: the 9-view 1920x1440 job (default); batch: 48 references "
                         "of uneven size in one captured batch, FP16 images, against the same "
                         "jobs run one by one in FP32; adjoint: cost + gradients of a 960x720 "
                         "depth map against 6 misaligned sources, and the 10-step polish; "
                         "reference: reference-mode PatchMatchStereo + RefineSubpixel of one "
                         "1280x960 view, the reference's CPU code on a 320x240 view as baseline "
                         "and bit-exactness check; "
                         "copy: a plain 1 GiB device copy (calibrates rocprofv3 byte counters)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1440)
    ap.add_argument("--iterations", type=int, default=5)
    ap.add_argument("--storage", default=None, choices=["f32", "f16", "bf16", "u8"],
                    help="image storage (default: u8 for --images u8, which holds them exactly; "
                         "f32 for --images float; f16 for --workload batch)")
    ap.add_argument("--images", default="u8", choices=["float", "u8"],
                    help="plane-mode input images: the renderer's images quantized to 8 bits "
                         "(what a camera and the reference deliver), or its float images")
    ap.add_argument("--view-phased", type=int, default=-1,
                    help="leading iterations run view-phased (one launch per source view); "
                         "-1: the library's automatic choice")
    ap.add_argument("--tile-schedule", type=int, default=-1,
                    help="workgroup tile schedule; -1 = autotuned with the per-phase occupancy "
                         "(during the warm-up step)")
    ap.add_argument("--check-size", default="320x240",
                    help="reference workload: size (WxH) of the view run by the reference's CPU "
                         "code as the baseline and compared with the GPU bit for bit")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target duration of the CPU-baseline sample (0 = skip)")
    return ap.parse_args()


def cpu_baseline(sc, iterations, target_s):
    """Reference PatchMatchStereo (fronto-parallel, 11x11, all 8 views, 5
    iterations x 4 directions) on a centred crop of view 0, all host cores."""
    import ctypes as C

    import oracle
    from recon_scanner_amd import _native as N
    if not oracle.reference_available():
        return None
    cores = len(os.sched_getaffinity(0))
    u8 = [sc.u8(i) for i in range(sc.num_views)]
    H, W = u8[0].shape

    def run(cw, ch, iters):
        x0, y0 = (W - cw) // 2, (H - ch) // 2
        crop = np.ascontiguousarray(u8[0][y0:y0 + ch, x0:x0 + cw])
        c0 = sc.cams[0]
        ref = N.Camera.make(c0.fx, c0.fy, c0.cx - x0, c0.cy - y0, cw, ch, c0.R_np(), c0.t_np())
        cfg = N.config_default()
        cfg.patch_size, cfg.iterations, cfg.num_threads = 11, iters, cores
        cfg.min_depth_m, cfg.max_depth_m = sc.min_depth, sc.max_depth
        cfg.num_views = 8
        poses = [ref] + list(sc.cams[1:])
        t = time.perf_counter()
        f, d, s, c = oracle.ref_patchmatch(crop, u8[1:], ref, poses, poses, list(range(1, 9)), cfg)
        dt = time.perf_counter() - t
        truth = sc.depths[0][y0:y0 + ch, x0:x0 + cw]
        m = np.isfinite(d) & (truth > 0)
        err = float(np.median(np.abs(d[m] - truth[m]))) if m.any() else float("nan")
        return dt, err

    # calibrate on a crop tall enough for every host thread (the reference
    # splits rows into min(threads, rows / 16) blocks), then size the sample
    # to ~target_s
    ch0 = min(H, max(120, 16 * cores))
    cw0 = min(W, max(160, ch0 * 4 // 3))
    dt, _ = run(cw0, ch0, 1)
    per_unit = dt / (cw0 * ch0 * 8 * 1)
    area = max(64 * 48, min(W * H, int(target_s / (per_unit * 8 * iterations))))
    ch = max(48, min(H, int((area * 3 / 4) ** 0.5)))
    cw = max(64, min(W, int(ch * 4 / 3)))
    dt, err = run(cw, ch, iterations)
    rate = cw * ch * 8 * iterations / dt / 1e9
    return {"value": rate, "unit": "Gpix*view*iter/s", "cores": cores, "kind": "reference",
            "sample": f"reference PatchMatchStereo (fronto-parallel, 11x11, 8 views, {iterations} "
                      f"iterations x 4 directions, num_threads={cores}) on a centred {cw}x{ch} "
                      f"crop of view 0; {dt:.1f} s",
            "median_depth_error_mm": err * 1000.0, "seconds": dt}


def base_record(args, dt, workload, **config):
    return {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": dt / args.steps * 1000.0, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "config": dict(workload=workload, **config)}


def median_rel_error(pm, sc, refs):
    errs = []
    for r in refs:
        d = pm.depth_map(r)["depth"]
        m = np.isfinite(d) & (sc.depths[r] > 0)
        errs.append(np.abs(d[m] - sc.depths[r][m]) / sc.depths[r][m])
    return float(np.median(np.concatenate(errs)))


def run_batch(args):
    """48 references (640x480 .. 1920x1440, 4-10 sources each) of a 64-camera
    orbit: one captured batch with FP16 images against the same jobs run one
    by one with FP32 images."""
    from recon_scanner_amd import plane, scenes
    storage = args.storage or "f16"
    t0 = time.time()
    sc, jobs = scenes.orbit_batch(num_refs=48, num_cams=64, seed=0)
    sys.stderr.write(f"[bench] orbit rendered in {time.time() - t0:.1f} s\n")
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=args.iterations,
                             min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0,
                             tile_schedule=args.tile_schedule)
    units = sum(sc.cams[r].width * sc.cams[r].height * len(s) for r, s in jobs) * args.iterations
    refs = [r for r, _ in jobs]
    # plain path: FP32 images, one plane_patchmatch per reference
    pm = plane.PlaneMVS(sc.cams, sc.images, storage="f32")

    def plain():
        for r, srcs in jobs:
            pm.patchmatch(r, srcs, p)
    for _ in range(args.warmup):
        plain()
    pm.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        plain()
    pm.synchronize()
    dt_plain = time.perf_counter() - t
    err_plain = median_rel_error(pm, sc, refs)
    sys.stderr.write(f"[bench] plain FP32 path: {dt_plain / args.steps * 1e3:.1f} ms per pass\n")
    # batched path
    pm = plane.PlaneMVS(sc.cams, sc.images, storage=storage)
    b = pm.batch(jobs, p)
    for _ in range(args.warmup):
        b.run()
    pm.synchronize()
    pm.timing(enable=True, reset=True)
    t = time.perf_counter()
    for _ in range(args.steps):
        b.run()
    pm.synchronize()
    dt = time.perf_counter() - t
    tim = pm.timing(enable=False, reset=True)
    err_b = median_rel_error(pm, sc, refs)
    b.close()
    mpx = sum(sc.cams[r].width * sc.cams[r].height for r in refs) / 1e6
    out = {"metric": "batched PatchMatch Gpix*view*iter/s", "value": units * args.steps / dt / 1e9,
           "unit": "Gpix*view*iter/s",
           "dtype": f"fp32 ({storage} images)",
           "data": "synthetic (64-camera orbit around a sphere and two boxes, analytic depth)"}
    out.update(base_record(args, dt, f"48 references of uneven size ({mpx:.1f} Mpx in total), "
                                     f"4-10 sources each, 11x11 stride 2, top-4, "
                                     f"{args.iterations} red-black iterations, one captured batch",
                           refs=len(refs), megapixels=mpx, storage=storage,
                           schedule=b.schedule))
    out.update({"depth_maps_per_s": len(refs) * args.steps / dt,
                "plain_fp32": {"value": units * args.steps / dt_plain / 1e9,
                               "depth_maps_per_s": len(refs) * args.steps / dt_plain,
                               "median_rel_depth_error": err_plain},
                "speedup_vs_plain_fp32": dt_plain / dt,
                "median_rel_depth_error": err_b,
                "batch_graph_ms": tim["batch"][0] / max(1, tim["batch"][1])})
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
    out.update(base_record(args, dt, "adjoint of a 960x720 depth map, 6 sources, 11x11 stride 2, "
                                     "top-4 (host copies included in ms_per_step)",
                           width=960, height=720, sources=6))
    out.update({"adjoint_kernel_ms": adj_ms, "polish_10_steps_ms": pol_ms})
    print(json.dumps(out))


def run_reference(args):
    """The drop-in path: reference-mode PatchMatchStereo + RefineSubpixel (the
    reference's algorithm, bit-exact) of one view of the sphere + planes scene
    against its 8 sources.  Throughput on the GPU at 1280x960 (or --width x
    --height); the reference's own CPU code (same config and num_threads, whose
    row-block RNG streams the GPU reproduces) on a 320x240 view (or
    --check-size) as the baseline, where the GPU's outputs are also checked
    against it bit for bit (the CPU takes ~20 s there on 16 cores, ~80 s at
    640x480)."""
    import oracle
    from recon_scanner_amd import mvs, scenes
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

    w, h = (args.width, args.height) if args.width != 1920 else (1280, 960)
    sc, u8, cfg = problem(w, h)
    L, ctx = mvs.lib(), mvs.context()

    def timers(enable, reset):
        ms, cnt = np.zeros(8), np.zeros(8, np.int64)
        mvs.check(L.recon_mvs_hip_timing(ctx, int(enable), int(reset), ms.ctypes.data,
                                         cnt.ctypes.data), "timing")
        return ms, cnt

    for _ in range(args.warmup):
        gpu_once(sc, u8, cfg)
    timers(True, True)
    t = time.perf_counter()
    for _ in range(args.steps):
        filled, refined, dm = gpu_once(sc, u8, cfg)
    dt = time.perf_counter() - t
    ms, cnt = timers(False, False)
    rec = base_record(args, dt, f"reference-mode PatchMatchStereo + RefineSubpixel, {w}x{h}, "
                      f"8 sources, 11x11, {args.iterations} iterations x 4 directions, "
                      f"num_threads={cores}", width=w, height=h, iterations=args.iterations)
    units = w * h * 8 * args.iterations
    rec.update(metric="reference-mode PatchMatchStereo px*view*iter/s (bit-exact with the reference)",
               value=units * args.steps / dt / 1e9, unit="Gpix*view*iter/s", dtype="fp64")
    truth = sc.depths[0]
    d = dm.depth.reshape(h, w)
    m = np.isfinite(d) & (truth > 0)
    rec["median_depth_error_mm"] = float(np.median(np.abs(d[m] - truth[m]))) * 1000.0
    # roofline of one directional pass (k_pm_pass and its k_pm_restart, timed
    # together by HIP events on the library's stream): nominal hypotheses per
    # pixel and pass H = 4
    # neighbours + random_search_count + restart_ratio, each over 8 views x
    # patch^2 sample pairs of REF_FLOP_PER_PAIR FP64 operations
    pass_ms = ms[0] / cnt[0] if cnt[0] else None
    hyps = 4 + cfg.random_search_count + cfg.restart_ratio
    flop = w * h * hyps * 8 * cfg.patch_size ** 2 * REF_FLOP_PER_PAIR
    achieved = flop / (pass_ms * 1e-3) / 1e12 if pass_ms else None
    traffic = None
    tf = os.path.join(ROOT, "profiles", "sweep_traffic.txt")
    if os.path.exists(tf):
        for line in open(tf):
            if line.startswith(f"traffic_bytes_per_launch reference {w}x{h} "):
                traffic = float(line.split()[-1])
    rec["roofline"] = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                       "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS if achieved else None,
                       "traffic": traffic,
                       "kernel": "one directional pass: k_pm_pass + k_pm_restart",
                       "mean_launch_ms": pass_ms, "flop_per_launch": flop,
                       "hypotheses_per_pixel": hyps}
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
    print(json.dumps(rec))


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
    out.update(base_record(args, dt, "1 GiB device-to-device copy"))
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "adjoint":
        return run_adjoint(args)
    if args.workload == "reference":
        return run_reference(args)
    if args.workload == "copy":
        return run_copy(args)
    # 8-bit images stored as 8 bits: the depth maps are bit-identical to FP32
    # storage (a GPU test holds this), at a quarter of the image bytes
    args.storage = args.storage or ("u8" if args.images == "u8" else "f32")
    from recon_scanner_amd import plane, scenes

    t0 = time.time()
    f = 2500.0 * args.width / 1920.0
    sc = scenes.sphere_planes(width=args.width, height=args.height, focal=f, n_views=9,
                              supersample=2)
    sys.stderr.write(f"[bench] scene rendered in {time.time() - t0:.1f} s\n")
    nv = sc.num_views
    views = list(range(nv))
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=args.iterations,
                             min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0,
                             tile_schedule=args.tile_schedule,
                             view_phased_iterations=args.view_phased)
    imgs = sc.images if args.images == "float" else [sc.u8(i) for i in range(nv)]
    pm = plane.PlaneMVS(sc.cams, imgs, storage=args.storage)

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
    sys.stderr.write(f"[bench] warmup done at {time.time() - t0:.1f} s\n")
    sys.stderr.flush()
    pm.timing(enable=True, reset=True)
    pm.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step()
    pm.synchronize()
    dt = time.perf_counter() - t
    tim = pm.timing(enable=False, reset=True)
    sys.stderr.write(f"[bench] {args.steps} steps in {dt:.3f} s; kernels {tim}\n")

    P = args.width * args.height
    V = nv - 1
    units = nv * P * V * args.iterations
    value = units * args.steps / dt / 1e9
    S = 36
    hve = nv * P * V * (1 + HYPS_PER_ITER * args.iterations)
    # roofline of the half-sweep (one k_plane_sweep launch, or the launches of a
    # view-phased half-sweep, timed together): algorithmic FLOPs / mean duration
    sweep_ms, sweep_n = tim["sweep"]
    flop_launch = (P / 2) * HYPS_PER_ITER * V * S * FLOP_PER_SAMPLE
    achieved = flop_launch / (sweep_ms / sweep_n * 1e-3) / 1e12 if sweep_n else None
    traffic = None
    tf = os.path.join(ROOT, "profiles", "sweep_traffic.txt")
    if os.path.exists(tf):
        for line in open(tf):
            if line.startswith(f"traffic_bytes_per_launch {args.width}x{args.height} {args.storage}"):
                traffic = float(line.split()[-1])

    # accuracy: median |Z - Z_true| over consistency-kept pixels of all maps
    errs = []
    kept = 0
    for r in views:
        d = pm.depth_map(r, drop_inconsistent=2)["depth"]
        m = np.isfinite(d) & (sc.depths[r] > 0)
        kept += int(m.sum())
        errs.append(np.abs(d[m] - sc.depths[r][m]))
    med = float(np.median(np.concatenate(errs))) * 1000.0

    cpu = None
    if args.cpu_seconds > 0:
        try:
            cpu = cpu_baseline(sc, args.iterations, args.cpu_seconds)
        except Exception as e:  # the baseline is reported, never required
            sys.stderr.write(f"[bench] cpu baseline failed: {e}\n")
    out = {
        "metric": "PatchMatch Gpix*view*iter/s (+ median depth error vs analytic truth)",
        "value": value,
        "unit": "Gpix*view*iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1000.0,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if (args.storage == "f32" or args.images == "u8") else f"fp32 ({args.storage} images)",
        "data": "synthetic (rendered sphere + 2 tilted planes, analytic depth"
                + (", 8-bit images)" if args.images == "u8" else ", float images)"),
        "config": {"workload": f"plane-mode PatchMatch, 9 views {args.width}x{args.height}, "
                               f"each view vs 8 sources, 11x11 stride 2, top-4, "
                               f"{args.iterations} red-black iterations + geometric-consistency "
                               f"filter over 9 depth maps",
                   "views": nv, "width": args.width, "height": args.height,
                   "iterations": args.iterations, "images": args.images,
                   "storage": args.storage, "view_phased_iterations": args.view_phased},
        "hypothesis_view_evals_per_s": hve * args.steps / dt,
        "median_depth_error_mm": med,
        "kept_pixels": kept,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": (achieved / PEAK_FP32_TFLOPS) if achieved else None,
                     "traffic": traffic,
                     "kernel": "half-sweep: k_plane_sweep (k_phase_prep + k_phase_view per source"
                               " + k_phase_final in view-phased iterations)",
                     "mean_launch_ms": sweep_ms / sweep_n if sweep_n else None,
                     "flop_per_launch": flop_launch},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
