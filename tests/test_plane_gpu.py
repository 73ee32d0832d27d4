"""GPU parity of plane mode (FP32 on the GPU) against the FP64 oracle.

Tolerances (stated per test):
  * per-view NCC of a hypothesis: |NCC_gpu - NCC_oracle| <= 1e-5 for every
    view valid on both sides (FP32 warp offsets and sums vs FP64 exact
    homography; measured maximum 3.7e-6);
  * aggregated cost: |c_gpu - c_oracle| <= 1e-5 * max(1, |c|) for every
    hypothesis with the same valid views;
  * random init planes: bit-exact (both sides generate them with the same FP32
    operations from the same counter RNG);
  * final depth after the red-black sweeps: within 0.1 % of the oracle's on
    >= 99 % of the pixels valid on both sides;
  * geometric-consistency support counts: bit-exact given the same depth maps.
"""
import sys

import numpy as np
import pytest

import oracle
from recon_scanner_amd import plane, scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def oparams(p):
    o = oracle.PlaneParams()
    for f, _ in oracle.PlaneParams._fields_:
        setattr(o, f, getattr(p, f))
    return o


def random_hyps(sc, n, rng, truth_frac=0.5):
    h, w = sc.depths[0].shape
    xs = rng.integers(0, w, n)
    ys = rng.integers(0, h, n)
    nrm = rng.normal(size=(n, 3))
    nrm[:, 2] = -np.abs(nrm[:, 2]) - 0.3
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.uniform(sc.min_depth, sc.max_depth, n)
    k = int(n * truth_frac)
    t = sc.depths[0][ys[:k], xs[:k]]
    d[:k] = np.where(t > 0, t * rng.uniform(0.995, 1.005, k), d[:k])
    nrm[:k] = [0, 0, -1.0]
    return np.concatenate([xs[:, None], ys[:, None], d[:, None], nrm], 1).astype(np.float32)


def compare_eval(sc, srcs, p, storage, n=4000, seed=0, min_valid=None):
    rng = np.random.default_rng(seed)
    imgs = [plane.round_storage(i, storage) for i in sc.images]
    pm = plane.PlaneMVS(sc.cams, sc.images, storage=storage)
    hyps = random_hyps(sc, n, rng)
    gc, gn, gv = pm.eval_costs(0, srcs, p, hyps)
    oc, on, ov = oracle.plane_eval(imgs, sc.cams, 0, srcs, oparams(p), hyps)
    both = np.isfinite(gv) & np.isfinite(ov)
    agree = np.mean(np.isfinite(gv) == np.isfinite(ov))
    dn = np.abs(gv[both] - ov[both])
    say(storage, "views valid both", int(both.sum()), "validity agreement", agree,
        "max |dNCC|", float(dn.max()), "p99.9", float(np.quantile(dn, 0.999)))
    d_all = np.where(both, np.abs(gv - ov), 0.0)
    for q in np.argsort(d_all.max(1))[::-1][:3]:
        v = int(np.argmax(d_all[q]))
        say("  worst: hyp", hyps[q].tolist(), "view", v, "gpu", float(gv[q, v]), "oracle", float(ov[q, v]),
            "nvalid", int(gn[q]), int(on[q]))
    assert agree > 0.999
    assert both.sum() > (n if min_valid is None else min_valid)
    assert dn.max() <= 1e-5
    fin = (gc < 1e29) & (oc < 1e29) & (gn == on)
    dc = np.abs(gc[fin] - oc[fin]) / np.maximum(1.0, np.abs(oc[fin]))
    say(storage, "aggregated cost: max relative |dc|", float(dc.max()))
    assert dc.max() <= 1e-5
    return gc, oc


def test_eval_sphere_arc_7x7_f32():
    say("eval sphere_arc 320x240 7x7 f32")
    sc = scenes.sphere_arc()
    p = plane.default_params(patch_size=7, window_stride=1, top_k=0, min_depth_m=0.25,
                             max_depth_m=0.45, value_scale=1.0)
    compare_eval(sc, [1, 2, 3, 4], p, "f32")


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_eval_sphere_planes_11x11_stride2_top4(storage):
    say("eval sphere_planes 480x360 11x11/2 top-4", storage)
    sc = scenes.sphere_planes(width=480, height=360, focal=625.0, supersample=1)
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, min_depth_m=0.25,
                             max_depth_m=0.60, value_scale=1.0)
    compare_eval(sc, list(range(1, 9)), p, storage, n=3000)


def test_init_planes_bit_exact_and_sweeps_match_oracle():
    say("plane patchmatch sphere_arc 320x240 7x7, 3 iterations")
    sc = scenes.sphere_arc()
    srcs = [1, 2, 3, 4]
    p0 = plane.default_params(patch_size=7, window_stride=1, top_k=0, iterations=0,
                              min_depth_m=0.25, max_depth_m=0.45, value_scale=1.0)
    pm = plane.PlaneMVS(sc.cams, sc.images)
    pm.patchmatch(0, srcs, p0)
    g0 = pm.depth_map(0)
    op, oc, _ = oracle.plane_patchmatch(sc.images, sc.cams, 0, srcs, oparams(p0))
    gpl = np.concatenate([g0["depth"][..., None], g0["normals"]], -1)
    fin = np.isfinite(g0["depth"])
    assert np.array_equal(gpl[..., 1:], op[..., 1:])           # normals bit-exact
    assert np.array_equal(g0["depth"][fin], op[..., 0][fin])    # depths bit-exact where scored
    p = p0.copy(iterations=3)
    pm.patchmatch(0, srcs, p)
    g = pm.depth_map(0)
    op, oc, _ = oracle.plane_patchmatch(sc.images, sc.cams, 0, srcs, oparams(p))
    od = np.where(oc < 1e29, op[..., 0], np.nan)
    both = np.isfinite(g["depth"]) & np.isfinite(od)
    rel = np.abs(g["depth"][both] - od[both]) / od[both]
    frac = np.mean(rel <= 1e-3)
    truth = sc.depths[0]
    m = both & (truth > 0)
    eg = np.median(np.abs(g["depth"][m] - truth[m]))
    eo = np.median(np.abs(od[m] - truth[m]))
    say("within 0.1%:", frac, "median |dZ| gpu", eg, "oracle", eo, "valid", both.sum())
    assert both.sum() > 0.95 * truth.size
    assert frac >= 0.99
    assert eg < 0.001


def test_geo_consistency_bit_exact():
    say("geo consistency 3 views 240x180")
    sc = scenes.sphere_planes(width=240, height=180, focal=312.5, n_views=3, supersample=1)
    p = plane.default_params(patch_size=11, window_stride=2, top_k=2, iterations=3,
                             min_depth_m=0.25, max_depth_m=0.60, value_scale=1.0)
    pm = plane.PlaneMVS(sc.cams, sc.images)
    views = [0, 1, 2]
    for r in views:
        pm.patchmatch(r, [v for v in views if v != r], p)
    pm.geo_consistency(views, 0.01)
    maps = [pm.depth_map(v, geo=True) for v in views]
    og = oracle.geo_consistency([m["depth"] for m in maps], sc.cams, 0.01)
    for m, o in zip(maps, og):
        assert np.array_equal(m["geo_support"], o)
    kept = np.mean(maps[0]["geo_support"][np.isfinite(maps[0]["depth"])] >= 2)
    say("kept fraction", kept)
    assert kept > 0.5


def test_refine_keeps_accuracy():
    say("plane refine sphere_arc")
    sc = scenes.sphere_arc()
    srcs = [1, 2, 3, 4]
    p = plane.default_params(patch_size=7, window_stride=1, top_k=0, iterations=3,
                             min_depth_m=0.25, max_depth_m=0.45, value_scale=1.0)
    pm = plane.PlaneMVS(sc.cams, sc.images)
    pm.patchmatch(0, srcs, p)
    a = pm.depth_map(0)
    pm.refine(0, srcs, p)
    b = pm.depth_map(0)
    truth = sc.depths[0]
    m = np.isfinite(a["depth"]) & np.isfinite(b["depth"]) & (truth > 0)
    ea = np.median(np.abs(a["depth"][m] - truth[m]))
    eb = np.median(np.abs(b["depth"][m] - truth[m]))
    say("median before", ea, "after", eb)
    assert np.all(b["cost"][m] <= a["cost"][m])
    assert eb <= ea * 1.05


@pytest.mark.parametrize("storage", ["f16", "u8"])
def test_batch_matches_per_view_runs_bit_exact(storage):
    """Batched persistent launch (uneven resolutions, 4-10 sources) ==
    per-view plane_patchmatch with the same parameters, over 3 graph replays
    with eager work between them, with the phases' binned form chosen by
    the tuner, forced on every phase and switched off.  FP16 storage of the float images, and
    8-bit storage of 8-bit images (what the stage uploads): the latter also
    equals per-view runs on the same images in FP32 storage bit for bit."""
    say("batch: orbit of 12 cameras, 6 refs, uneven sizes,", storage)
    sc, jobs = scenes.orbit_batch(num_refs=6, num_cams=12, seed=3, min_res=(160, 120),
                                  max_res=(320, 240), src_range=(4, 6))
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=2,
                             min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0,
                             tile_schedule=1)
    imgs = sc.images if storage == "f16" else [sc.u8(i) for i in range(sc.num_views)]
    pm = plane.PlaneMVS(sc.cams, imgs, storage=storage)
    singles = {}
    for r, srcs in jobs:
        pm.patchmatch(r, srcs, p.copy(binned_phases=0))
        singles[r] = pm.depth_map(r)
    if storage == "u8":
        pf = plane.PlaneMVS(sc.cams, imgs, storage="f32")
        for r, srcs in jobs:
            pf.patchmatch(r, srcs, p)
            f = pf.depth_map(r)
            for k in ("depth", "normals", "cost", "support"):
                assert f[k].tobytes() == singles[r][k].tobytes(), (r, k)
    # tuned plan (binned_phases = -1), every phase binned, none binned
    for binned in (-1, 1 + 2 * p.iterations, 0):
        b = pm.batch(jobs, p.copy(binned_phases=binned))
        for rep in range(3):   # graph replays, eager work between them
            b.run()
            pm.synchronize()
            for r, _ in jobs:
                g = pm.depth_map(r)
                for k in ("depth", "normals", "cost", "support"):
                    assert g[k].tobytes() == singles[r][k].tobytes(), (binned, rep, r, k)
        b.close()


def test_autotune_picks_a_compiled_schedule():
    say("autotune")
    sc = scenes.sphere_planes(width=480, height=360, focal=625.0, supersample=1)
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=1,
                             min_depth_m=0.25, max_depth_m=0.60, value_scale=1.0, tile_schedule=-1)
    pm = plane.PlaneMVS(sc.cams, sc.images)
    best, ms = pm.autotune(0, list(range(1, 9)), p)
    say("schedule ms", ms, "best", best)
    assert best in (0, 1, 2) and ms[best] > 0 and ms[best] == min(m for m in ms if m > 0)
    pm.patchmatch(0, list(range(1, 9)), p)     # -1 resolves through the cache
    assert np.isfinite(pm.depth_map(0)["depth"]).mean() > 0.5


def test_occupancy_caps_do_not_change_results():
    """The per-phase occupancy caps change only how many workgroups share a
    CU: depth maps are bit-identical under every setting."""
    say("occupancy caps: -1, 1, 2, 3 on sphere_planes 480x360, 2 iterations")
    sc = scenes.sphere_planes(width=480, height=360, focal=625.0, supersample=1)
    srcs = list(range(1, 9))
    pm = plane.PlaneMVS(sc.cams, sc.images)
    ref = None
    for cap in (-1, 1, 2, 3):
        p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=2,
                                 min_depth_m=0.25, max_depth_m=0.60, value_scale=1.0,
                                 tile_schedule=1, occupancy_cap=cap, capped_iterations=2)
        pm.patchmatch(0, srcs, p)
        g = pm.depth_map(0)
        if ref is None:
            ref = g
            continue
        for k in ("depth", "normals", "cost", "support"):
            assert np.array_equal(np.nan_to_num(g[k], nan=-1), np.nan_to_num(ref[k], nan=-1)), (cap, k)


def test_bgra_frames_gray_conversion_bit_exact():
    """B,G,R,A camera frames: the GPU conversion equals BgraToGray
    (orchestrator.cpp:53-67, restated in the oracle), standalone and fused into
    the upload; depth maps from BGRA views equal those from the gray views."""
    say("BGRA8 upload, sphere_arc 160x120, padded rows")
    sc = scenes.sphere_arc(width=160, height=120, focal=250.0)
    rng = np.random.default_rng(5)
    frames, grays = [], []
    for i in range(sc.num_views):
        g = sc.u8(i).astype(np.int32)
        tint = rng.integers(-20, 21, size=(3,))
        bgr = np.clip(g[..., None] + tint[None, None, :] + rng.integers(-3, 4, g.shape + (3,)), 0, 255)
        buf = np.zeros((120, 160 * 4 + 36), np.uint8)            # rows padded to 676 bytes
        view = buf[:, :640].reshape(120, 160, 4)
        view[..., :3] = bgr
        view[..., 3] = 255
        frames.append(view)
        grays.append(oracle.bgra_to_gray(view))
    assert np.array_equal(plane.bgra_to_gray(frames[1]), grays[1])
    p = plane.default_params(patch_size=7, window_stride=1, top_k=0, iterations=2,
                             min_depth_m=0.25, max_depth_m=0.45, value_scale=255.0, tile_schedule=0)
    a = plane.PlaneMVS(sc.cams, frames)
    a.patchmatch(0, [1, 2, 3, 4], p)
    da = a.depth_map(0)
    b = plane.PlaneMVS(sc.cams, grays)
    b.patchmatch(0, [1, 2, 3, 4], p)
    db = b.depth_map(0)
    for k in ("depth", "normals", "cost"):
        assert np.array_equal(np.nan_to_num(da[k], nan=-1), np.nan_to_num(db[k], nan=-1)), k
    assert np.isfinite(da["depth"]).mean() > 0.9
    # BGRA frames into 8-bit storage (the 11x11 stride-2 window), against the
    # gray views in FP32 storage
    p2 = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=2,
                              min_depth_m=0.25, max_depth_m=0.45, value_scale=1.0, tile_schedule=0)
    c = plane.PlaneMVS(sc.cams, frames, storage="u8")
    c.patchmatch(0, [1, 2, 3, 4], p2)
    dc = c.depth_map(0)
    d = plane.PlaneMVS(sc.cams, grays, storage="f32")
    d.patchmatch(0, [1, 2, 3, 4], p2)
    dd = d.depth_map(0)
    for k in ("depth", "normals", "cost", "support"):
        assert dc[k].tobytes() == dd[k].tobytes(), k


def test_seeded_init_bit_exact_and_converges_faster():
    """Sparse-point seeding (patchmatch_stereo.cpp:499-533): init planes
    bit-exact with the oracle's (seeded cells fronto-parallel at the FP64 cell
    mean, the rest the random init), sweeps from there within 0.1 % of the
    oracle's on >= 99 % of the pixels, and seeds from 2 %-noisy sparse depths
    reach after one iteration an accuracy the random init does not."""
    say("plane patchmatch seeded sphere_arc 320x240 7x7")
    sc = scenes.sphere_arc()
    srcs = [1, 2, 3, 4]
    h, w = sc.depths[0].shape
    rng = np.random.default_rng(4)
    n = 300
    us, vs = rng.uniform(0, w, n), rng.uniform(0, h, n)
    t = sc.depths[0][vs.astype(int), us.astype(int)]
    seeds = np.stack([us, vs, t * rng.uniform(0.98, 1.02, n)], 1)
    seeds = seeds[t > 0]
    p0 = plane.default_params(patch_size=7, window_stride=1, top_k=0, iterations=0,
                              min_depth_m=0.25, max_depth_m=0.45, value_scale=1.0)
    pm = plane.PlaneMVS(sc.cams, sc.images)
    pm.patchmatch(0, srcs, p0, seed_points=seeds)
    g0 = pm.depth_map(0)
    init = oracle.seeded_init_planes(w, h, seeds)
    op, oc, _ = oracle.plane_patchmatch(sc.images, sc.cams, 0, srcs, oparams(p0), init_planes=init)
    fin = np.isfinite(g0["depth"])
    assert np.array_equal(g0["normals"], op[..., 1:])
    assert np.array_equal(g0["depth"][fin], op[..., 0][fin])
    seeded = np.isfinite(init[..., 0])
    assert seeded.mean() > 0.9 and np.all(g0["normals"][seeded] == [0.0, 0.0, -1.0])
    p = p0.copy(iterations=2)
    pm.patchmatch(0, srcs, p, seed_points=seeds)
    g = pm.depth_map(0)
    op, oc, _ = oracle.plane_patchmatch(sc.images, sc.cams, 0, srcs, oparams(p), init_planes=init)
    od = np.where(oc < 1e29, op[..., 0], np.nan)
    both = np.isfinite(g["depth"]) & np.isfinite(od)
    rel = np.abs(g["depth"][both] - od[both]) / od[both]
    say("seeded 2 it: within 0.1%:", np.mean(rel <= 1e-3), "valid", both.sum())
    assert np.mean(rel <= 1e-3) >= 0.99
    # after one iteration, fraction of pixels within 0.5 mm of the truth
    truth = sc.depths[0]

    def good(d):
        m = truth > 0
        return float(np.mean(np.abs(np.nan_to_num(d[m], nan=0.0) - truth[m]) < 5e-4))

    p1 = p0.copy(iterations=1)
    pm.patchmatch(0, srcs, p1, seed_points=seeds)
    gs = good(pm.depth_map(0)["depth"])
    pm.patchmatch(0, srcs, p1)
    dr = pm.depth_map(0)["depth"]
    gr = good(dr)
    say("within 0.5 mm after 1 iteration: seeded", gs, "random", gr)
    assert gs > gr
    # an empty seed list is the unseeded call
    pm.patchmatch(0, srcs, p1, seed_points=np.zeros((0, 3)))
    assert np.array_equal(pm.depth_map(0)["depth"], dr, equal_nan=True)


@pytest.mark.parametrize("storage,top_k", [("f32", 4), ("f16", 4), ("f32", 0), ("u8", 4)])
def test_view_phased_sweeps_bit_identical(storage, top_k):
    """The view-phased half-sweep (one launch per source view with its own
    tile shape, candidate state in memory between launches, early exit
    against the cost from the start of the sweep) gives the fused sweep's
    planes, costs and supports bit for bit, for top-k and all-views
    aggregation; and every tile schedule gives the same results."""
    say("view-phased vs fused sweeps, sphere_planes 480x360 11x11/2", storage, "top_k", top_k)
    sc = scenes.sphere_planes(width=480, height=360, focal=625.0, supersample=1)
    srcs = list(range(1, 9))
    imgs = [sc.u8(i) for i in range(sc.num_views)] if storage == "u8" else sc.images
    pm = plane.PlaneMVS(sc.cams, imgs, storage=storage)
    p = plane.default_params(patch_size=11, window_stride=2, top_k=top_k, iterations=3,
                             min_depth_m=0.25, max_depth_m=0.60, value_scale=1.0, tile_schedule=2)
    outs = []
    for phased, sched in ((0, 2), (2, 2), (3, 2), (0, 0), (0, 1), (1, 1)):
        pm.patchmatch(0, srcs, p.copy(view_phased_iterations=phased, tile_schedule=sched))
        outs.append(pm.depth_map(0))
    for o in outs[1:]:
        for k in ("depth", "normals", "cost", "support"):
            assert np.array_equal(o[k], outs[0][k], equal_nan=True), k
    fin = np.isfinite(outs[0]["depth"])
    assert fin.mean() > 0.9
    # seeded init (phased init reproduces the seeding too)
    rng = np.random.default_rng(2)
    h, w = sc.depths[0].shape
    us, vs = rng.uniform(0, w, 400), rng.uniform(0, h, 400)
    t = sc.depths[0][vs.astype(int), us.astype(int)]
    seeds = np.stack([us, vs, t], 1)[t > 0]
    seeded = []
    for phased in (0, 1):
        pm.patchmatch(0, srcs, p.copy(view_phased_iterations=phased, iterations=1), seed_points=seeds)
        seeded.append(pm.depth_map(0))
    for k in ("depth", "normals", "cost", "support"):
        assert np.array_equal(seeded[1][k], seeded[0][k], equal_nan=True), k


@pytest.mark.parametrize("storage,top_k,reach", [("f32", 4, 4), ("f16", 4, 4), ("u8", 4, 4), ("f16", 0, 4),
                                                 ("bf16", 1, 1), ("f16", 2, 8)])
def test_binned_phases_bit_identical(storage, top_k, reach):
    """Binned init and half-sweeps (each tile's hypotheses sorted by inverse
    depth in LDS and evaluated bin by bin; early exit against each pixel's
    best cost so far) give the fused kernels' planes, costs and
    supports bit for bit: on every phase, on the leading two, under every
    tile schedule, for top-k and all-views aggregation, with random and
    seeded init, and under the tuned plan."""
    say("binned vs fused phases, sphere_planes 480x360 11x11/2", storage, "top_k", top_k, "reach", reach)
    sc = scenes.sphere_planes(width=480, height=360, focal=625.0, supersample=1)
    srcs = list(range(1, 9))
    imgs = [sc.u8(i) for i in range(sc.num_views)] if storage == "u8" else sc.images
    pm = plane.PlaneMVS(sc.cams, imgs, storage=storage)
    p = plane.default_params(patch_size=11, window_stride=2, top_k=top_k, iterations=3,
                             min_depth_m=0.25, max_depth_m=0.60, value_scale=1.0,
                             view_phased_iterations=0, occupancy_cap=-1, propagation_reach=reach)
    outs = []
    for binned, sched in ((0, 0), (7, 0), (2, 0), (7, 1), (7, 2), (0, 2)):
        pm.patchmatch(0, srcs, p.copy(binned_phases=binned, tile_schedule=sched))
        outs.append(pm.depth_map(0))
    # tuned schedule and per-phase configuration (binned among the trials)
    pm.patchmatch(0, srcs, p.copy(binned_phases=-1, tile_schedule=-1, occupancy_cap=0))
    outs.append(pm.depth_map(0))
    for o in outs[1:]:
        for k in ("depth", "normals", "cost", "support"):
            assert np.array_equal(o[k], outs[0][k], equal_nan=True), k
    assert np.isfinite(outs[0]["depth"]).mean() > 0.9
    rng = np.random.default_rng(5)
    h, w = sc.depths[0].shape
    us, vs = rng.uniform(0, w, 400), rng.uniform(0, h, 400)
    t = sc.depths[0][vs.astype(int), us.astype(int)]
    seeds = np.stack([us, vs, t], 1)[t > 0]
    seeded = []
    for binned in (0, 3):
        pm.patchmatch(0, srcs, p.copy(binned_phases=binned, iterations=1, tile_schedule=1),
                      seed_points=seeds)
        seeded.append(pm.depth_map(0))
    for k in ("depth", "normals", "cost", "support"):
        assert np.array_equal(seeded[1][k], seeded[0][k], equal_nan=True), k


def test_binned_ragged_sizes_bit_identical():
    """Binned and fused forms agree bit for bit on images that no tile
    schedule divides, narrower or shorter than a tile (pixels outside the
    image take no entries), per view under every schedule and in one batch
    of all the sizes; and the batch equals the per-view runs."""
    sizes = [(101, 37), (47, 211), (150, 9), (64, 8)]
    say("binned vs fused on ragged sizes", sizes)
    cams, imgs, jobs = [], [], []
    for w, h in sizes:
        sc = scenes.sphere_planes(width=w, height=h, focal=2500.0 * w / 1920.0 * 4.0, supersample=1,
                                  n_views=5)
        base = len(cams)
        cams += list(sc.cams)
        imgs += list(sc.images)
        jobs.append((base, [base + k for k in range(1, 5)]))
    pm = plane.PlaneMVS(cams, imgs, storage="f16")
    p = plane.default_params(patch_size=11, window_stride=2, top_k=2, iterations=2,
                             min_depth_m=0.25, max_depth_m=0.60, value_scale=1.0,
                             view_phased_iterations=0, occupancy_cap=-1)
    singles = {}
    for r, srcs in jobs:
        outs = []
        for binned, sched in ((0, 0), (5, 0), (5, 1), (5, 2), (0, 2)):
            pm.patchmatch(r, srcs, p.copy(binned_phases=binned, tile_schedule=sched))
            outs.append(pm.depth_map(r))
        for o in outs[1:]:
            for k in ("depth", "normals", "cost", "support"):
                assert np.array_equal(o[k], outs[0][k], equal_nan=True), (r, k)
        singles[r] = outs[0]
    assert any(np.isfinite(o["depth"]).mean() > 0.5 for o in singles.values())
    for binned in (5, 0):
        b = pm.batch(jobs, p.copy(binned_phases=binned, tile_schedule=1))
        b.run()
        pm.synchronize()
        for r, _ in jobs:
            g = pm.depth_map(r)
            for k in ("depth", "normals", "cost", "support"):
                assert g[k].tobytes() == singles[r][k].tobytes(), (binned, r, k)
        b.close()


def test_u8_images_half_storage_bit_identical():
    """8-bit images stored as FP16, BF16 or 8 bits (all hold the integers
    0..255 exactly) give the FP32-storage results bit for bit: planes, costs,
    supports, confidences, after refinement and polish, and the consistency
    filter over them.  8-bit views are stored as 8 bits by default (the
    depth-map bench and the stage) on this ground."""
    say("u8 images: f32 vs f16 vs bf16 storage, sphere_planes 480x360 11x11/2")
    sc = scenes.sphere_planes(width=480, height=360, focal=625.0, supersample=1, n_views=5)
    u8 = [sc.u8(i) for i in range(sc.num_views)]
    views = list(range(sc.num_views))
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=3,
                             min_depth_m=0.25, max_depth_m=0.60, value_scale=1.0)
    res = {}
    for storage in ("f32", "f16", "bf16", "u8"):
        pm = plane.PlaneMVS(sc.cams, u8, storage=storage)
        for r in views:
            srcs = [v for v in views if v != r]
            pm.patchmatch(r, srcs, p)
            if r == 0:  # the sub-pixel refinement and the gradient polish too
                pm.refine(r, srcs, p)
                pm.polish(r, srcs, p, steps=5)
        pm.geo_consistency(views, 0.01)
        res[storage] = [pm.depth_map(r, geo=True) for r in views]
    for storage in ("f16", "bf16", "u8"):
        for a, b in zip(res["f32"], res[storage]):
            for k in a:
                assert a[k].tobytes() == b[k].tobytes(), (storage, k)
    assert np.isfinite(res["f32"][0]["depth"]).mean() > 0.9
    # 8-bit storage refuses views that are not 8-bit
    with pytest.raises(Exception):
        plane.PlaneMVS(sc.cams, sc.images, storage="u8")


def test_tuned_plan_bit_identical():
    """tile_schedule = -1 with occupancy_cap = 0 (the tuned schedule and
    per-phase occupancy caps) on 8-bit images in 8-bit storage (the
    depth-map bench's configuration) and in FP16: the results equal every
    fixed schedule's bit for bit."""
    say("tuned plan vs fixed schedules, sphere_planes 480x360 u8 images, 3 iterations")
    sc = scenes.sphere_planes(width=480, height=360, focal=625.0, supersample=1)
    u8 = [sc.u8(i) for i in range(sc.num_views)]
    srcs = list(range(1, 9))
    for storage in ("u8", "f16"):
        _tuned_plan_case(sc, u8, srcs, storage)


def _tuned_plan_case(sc, u8, srcs, storage):
    pm = plane.PlaneMVS(sc.cams, u8, storage=storage)
    base = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=3,
                                min_depth_m=0.25, max_depth_m=0.60, value_scale=1.0)
    outs = []
    for sched in (-1, 0, 1, 2, -1):
        pm.patchmatch(0, srcs, base.copy(tile_schedule=sched, occupancy_cap=0))
        outs.append(pm.depth_map(0))
    for o in outs[1:]:
        for k in ("depth", "normals", "cost", "support"):
            assert np.array_equal(o[k], outs[0][k], equal_nan=True), k
    assert np.isfinite(outs[0]["depth"]).mean() > 0.9


def orbit_16_sources(res=(256, 192)):
    """Reference 0 of a 64-camera orbit with its 8 neighbours on each side as
    sources: 16, the most a reference takes (baselines up to 45 degrees)."""
    sc, _ = scenes.orbit_batch(num_refs=1, num_cams=64, seed=7, min_res=res, max_res=res, src_range=(4, 4))
    return sc, [d % 64 for d in range(-8, 9) if d != 0]


@pytest.mark.parametrize("top_k", [8, 5])
def test_eval_16_sources_top_k_above_4_matches_oracle(top_k):
    """Per-view NCC and aggregated cost with 16 sources and top-k above 4
    (the 8-wide insertion network), against the FP64 oracle at the bars
    above."""
    say("eval orbit 256x192 11x11/2, 16 sources, top_k", top_k)
    sc, srcs = orbit_16_sources()
    p = plane.default_params(patch_size=11, window_stride=2, top_k=top_k, min_depth_m=sc.min_depth,
                             max_depth_m=sc.max_depth, value_scale=1.0)
    compare_eval(sc, srcs, p, "f32", n=2000)


def test_binned_16_sources_top8_bit_identical():
    """Binned and fused forms agree bit for bit with 16 sources and top-8."""
    say("binned vs fused, orbit 256x192 11x11/2, 16 sources, top-8")
    sc, srcs = orbit_16_sources()
    pm = plane.PlaneMVS(sc.cams, sc.images, storage="f16")
    p = plane.default_params(patch_size=11, window_stride=2, top_k=8, iterations=2,
                             min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0,
                             view_phased_iterations=0, occupancy_cap=-1)
    outs = []
    for binned, sched in ((0, 0), (5, 0), (5, 1), (5, 2)):
        pm.patchmatch(0, srcs, p.copy(binned_phases=binned, tile_schedule=sched))
        outs.append(pm.depth_map(0))
    pm.patchmatch(0, srcs, p.copy(binned_phases=-1, tile_schedule=-1, occupancy_cap=0))
    outs.append(pm.depth_map(0))
    for o in outs[1:]:
        for k in ("depth", "normals", "cost", "support"):
            assert np.array_equal(o[k], outs[0][k], equal_nan=True), k
    assert np.isfinite(outs[0]["depth"]).mean() > 0.5


def test_batch_16_sources_top8_matches_per_view():
    """A batch whose jobs take 16 sources each (top-8) equals per-view runs
    bit for bit, over two graph replays."""
    say("batch, orbit 64 cameras 192x144, 3 refs x 16 sources, top-8")
    sc, _ = scenes.orbit_batch(num_refs=1, num_cams=64, seed=9, min_res=(192, 144), max_res=(192, 144),
                               src_range=(4, 4))
    jobs = [(r, [(r + d) % 64 for d in range(-8, 9) if d != 0]) for r in (0, 21, 42)]
    p = plane.default_params(patch_size=11, window_stride=2, top_k=8, iterations=2,
                             min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0,
                             tile_schedule=1)
    pm = plane.PlaneMVS(sc.cams, sc.images, storage="f16")
    singles = {}
    for r, srcs in jobs:
        pm.patchmatch(r, srcs, p.copy(binned_phases=0))
        singles[r] = pm.depth_map(r)
    b = pm.batch(jobs, p.copy(binned_phases=-1))
    for rep in range(2):
        b.run()
        pm.synchronize()
        for r, _ in jobs:
            g = pm.depth_map(r)
            for k in ("depth", "normals", "cost", "support"):
                assert g[k].tobytes() == singles[r][k].tobytes(), (rep, r, k)
    b.close()


@pytest.mark.parametrize("nsrc,top_k", [(1, 1), (2, 1), (2, 0)])
def test_few_sources_match_oracle_and_binned_bit_identical(nsrc, top_k):
    """One or two sources (all-views mean or best-1): costs against the FP64
    oracle at the bars above, and binned and fused sweeps bit for bit."""
    say("few sources, sphere_planes 480x360 11x11/2, nsrc", nsrc, "top_k", top_k)
    sc = scenes.sphere_planes(width=480, height=360, focal=625.0, supersample=1)
    srcs = [4, 5][:nsrc]
    p = plane.default_params(patch_size=11, window_stride=2, top_k=top_k, iterations=2,
                             min_depth_m=0.25, max_depth_m=0.60, value_scale=1.0,
                             view_phased_iterations=0, occupancy_cap=-1)
    compare_eval(sc, srcs, p, "f32", n=3000, min_valid=1500 * nsrc)
    pm = plane.PlaneMVS(sc.cams, sc.images, storage="f16")
    outs = []
    for binned, sched in ((0, 0), (5, 0), (5, 2)):
        pm.patchmatch(0, srcs, p.copy(binned_phases=binned, tile_schedule=sched))
        outs.append(pm.depth_map(0))
    for o in outs[1:]:
        for k in ("depth", "normals", "cost", "support"):
            assert np.array_equal(o[k], outs[0][k], equal_nan=True), k
    assert np.isfinite(outs[0]["depth"]).mean() > 0.5
