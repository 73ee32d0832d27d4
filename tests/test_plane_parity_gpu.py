"""GPU parity of plane mode at the shapes the benchmarks run, against the FP64
oracle (oracle/pm_oracle_plane.c).

Bars (stated per test):
  * random init planes: bit-exact (both sides generate them with the same
    FP32 operations from the same counter RNG);
  * depth after the red-black sweeps (or after refinement): within 0.1 % of
    the oracle's on >= 99 % of the pixels valid on both sides;
  * refinement: the cost of the GPU's refined plane within 1e-5 (relative to
    max(1, |c|)) of the oracle's refined cost on >= 99 % of those pixels;
  * batch: bit-identical to per-view runs; the full 48-reference batch within
    the accuracy of the per-view path against the analytic truth.
"""
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


def depth_agreement(g, op, oc):
    """Fraction of pixels valid on both sides whose depths agree within 0.1 %."""
    od = np.where(oc < 1e29, op[..., 0], np.nan)
    both = np.isfinite(g["depth"]) & np.isfinite(od)
    rel = np.abs(g["depth"][both] - od[both]) / od[both]
    return float(np.mean(rel <= 1e-3)), int(both.sum()), both


def test_sweeps_match_oracle_at_bench_window():
    """The bench's plane-mode configuration -- 11x11 window with stride 2,
    top-4 of 8 sources, 8-bit images in 8-bit storage -- on the sphere +
    tilted planes scene at 480x360, 3 red-black iterations: init planes
    bit-exact, final depths within 0.1 % of the oracle's on >= 99 % of the
    pixels valid on both sides."""
    say("sweeps vs oracle: sphere_planes 480x360, 11x11/2, top-4, 8 sources, u8 storage, 3 iterations")
    sc = scenes.sphere_planes(width=480, height=360, focal=625.0, supersample=1)
    u8 = [sc.u8(i) for i in range(sc.num_views)]
    srcs = list(range(1, 9))
    p0 = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=0,
                              min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=255.0,
                              tile_schedule=0)
    pm = plane.PlaneMVS(sc.cams, u8, storage="u8")
    imgs = [i.astype(np.float32) for i in u8]
    pm.patchmatch(0, srcs, p0)
    g0 = pm.depth_map(0)
    op, oc, _ = oracle.plane_patchmatch(imgs, sc.cams, 0, srcs, oparams(p0))
    fin = np.isfinite(g0["depth"])
    assert np.array_equal(g0["normals"], op[..., 1:])
    assert np.array_equal(g0["depth"][fin], op[..., 0][fin])
    p = p0.copy(iterations=3, tile_schedule=-1)
    pm.patchmatch(0, srcs, p)
    g = pm.depth_map(0)
    op, oc, _ = oracle.plane_patchmatch(imgs, sc.cams, 0, srcs, oparams(p))
    frac, nboth, both = depth_agreement(g, op, oc)
    truth = sc.depths[0]
    m = both & (truth > 0)
    eg = np.median(np.abs(g["depth"][m] - truth[m]))
    eo = np.median(np.abs(op[..., 0][m] - truth[m]))
    say("within 0.1%:", frac, "valid both", nboth, "median |dZ| gpu", eg, "oracle", eo)
    assert nboth > 0.9 * truth.size
    assert frac >= 0.99
    assert eg < 0.001


@pytest.mark.parametrize("case", ["sphere_arc_7x7", "sphere_planes_11x11s2_top4"])
def test_refine_matches_oracle(case):
    """Plane-mode sub-pixel refinement (k_plane_refine) against the oracle's
    FP64 restatement of the same search (orc_plane_refine), started from the
    same planes (the GPU's PatchMatch result): refined depths within 0.1 % on
    >= 99 % of the pixels, refined costs within 1e-5 on >= 99 %, and the GPU's
    refined cost never above its starting cost."""
    if case == "sphere_arc_7x7":
        sc = scenes.sphere_arc()
        srcs = [1, 2, 3, 4]
        p = plane.default_params(patch_size=7, window_stride=1, top_k=0, iterations=3,
                                 min_depth_m=0.25, max_depth_m=0.45, value_scale=1.0)
        imgs = sc.images
        storage = "f32"
    else:
        sc = scenes.sphere_planes(width=480, height=360, focal=625.0, supersample=1)
        srcs = list(range(1, 9))
        p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=3,
                                 min_depth_m=sc.min_depth, max_depth_m=sc.max_depth,
                                 value_scale=255.0)
        imgs = [sc.u8(i) for i in range(sc.num_views)]
        storage = "u8"
    say("refine vs oracle", case)
    pm = plane.PlaneMVS(sc.cams, imgs, storage=storage)
    pm.patchmatch(0, srcs, p)
    a = pm.depth_map(0)
    pm.refine(0, srcs, p)
    b = pm.depth_map(0)
    fimgs = [np.asarray(i, np.float32) for i in imgs]
    valid = np.isfinite(a["depth"])
    planes = np.concatenate([np.where(valid, a["depth"], 1.0)[..., None], a["normals"]], -1)
    h, w = valid.shape
    yy, xx = np.nonzero(valid)
    hyps = np.concatenate([xx[:, None], yy[:, None], planes[valid]], 1).astype(np.float32)
    c0, nv0, _ = oracle.plane_eval(fimgs, sc.cams, 0, srcs, oparams(p), hyps)
    cost = np.full((h, w), 1e30)
    sup = np.zeros((h, w), np.uint8)
    cost[valid] = c0
    sup[valid] = nv0
    rp, rc, _ = oracle.plane_refine(fimgs, sc.cams, 0, srcs, oparams(p), planes, cost, sup)
    both = np.isfinite(b["depth"]) & (rc < 1e29) & valid
    rel = np.abs(b["depth"][both] - rp[..., 0][both]) / rp[..., 0][both]
    dc = np.abs(b["cost"][both] - rc[both]) / np.maximum(1.0, np.abs(rc[both]))
    moved = np.mean(b["depth"][both] != a["depth"][both])
    say("depth within 0.1%:", np.mean(rel <= 1e-3), "cost within 1e-5:", np.mean(dc <= 1e-5),
        "max |dc|", float(dc.max()), "moved", moved, "pixels", int(both.sum()))
    assert both.sum() > 0.9 * valid.sum()
    assert moved > 0.5
    assert np.mean(rel <= 1e-3) >= 0.99
    assert np.mean(dc <= 1e-5) >= 0.99
    fin = np.isfinite(a["depth"]) & np.isfinite(b["depth"])
    assert np.all(b["cost"][fin] <= a["cost"][fin])


ORBIT_SUBSET = dict(num_refs=3, num_cams=64, seed=7, min_res=(640, 480), max_res=(1280, 960))


@pytest.fixture(scope="module")
def orbit_subset():
    sc, jobs = scenes.orbit_batch(**ORBIT_SUBSET)
    return sc, jobs


@pytest.mark.parametrize("job", [0, 1, 2])
def test_batch_orbit_subset_matches_oracle(orbit_subset, job):
    """The batched workload's own scene and shapes (64-camera orbit around a
    sphere and two boxes; references of 960x720 to 1248x936 with 7-10
    sources; FP16 images), one captured batch of the 3 references, 3
    iterations: each reference's init planes bit-exact with the oracle's (on
    the FP16-rounded images), its final depths within 0.1 % of the oracle's
    on >= 99 % of the pixels valid on both sides."""
    sc, jobs = orbit_subset
    r, srcs = jobs[job]
    say("batch vs oracle: orbit subset job", job, "ref", r, sc.cams[r].width, "x", sc.cams[r].height,
        "sources", len(srcs))
    p0 = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=0,
                              min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0,
                              tile_schedule=0)
    pm = plane.PlaneMVS(sc.cams, sc.images, storage="f16")
    imgs = [plane.round_storage(i, "f16") for i in sc.images]
    b = pm.batch(jobs, p0)
    b.run()
    pm.synchronize()
    b.close()
    g0 = pm.depth_map(r)
    op, oc, _ = oracle.plane_patchmatch(imgs, sc.cams, r, srcs, oparams(p0))
    fin = np.isfinite(g0["depth"])
    assert np.array_equal(g0["normals"], op[..., 1:])
    assert np.array_equal(g0["depth"][fin], op[..., 0][fin])
    p = p0.copy(iterations=3)
    b = pm.batch(jobs, p)
    b.run()
    pm.synchronize()
    b.close()
    g = pm.depth_map(r)
    op, oc, _ = oracle.plane_patchmatch(imgs, sc.cams, r, srcs, oparams(p))
    frac, nboth, _ = depth_agreement(g, op, oc)
    say("within 0.1%:", frac, "valid both", nboth)
    assert nboth > 0.5 * g["depth"].size
    assert frac >= 0.99


def test_batch_full_orbit_48_references():
    """The full batched workload (48 references of 640x480 to 1920x1440, 4-10
    sources each, FP16 images, 5 iterations, one graph): every job's map is
    filled, the median relative depth error against the analytic truth is
    that of the per-view FP32 path within 5 %, and two jobs equal their
    per-view FP16 runs bit for bit."""
    sc, jobs = scenes.orbit_batch(num_refs=48, num_cams=64, seed=0)
    say("full orbit batch: 48 refs,", sum(sc.cams[r].width * sc.cams[r].height for r, _ in jobs) / 1e6,
        "Mpx")
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=5,
                             min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0,
                             tile_schedule=-1)
    pm = plane.PlaneMVS(sc.cams, sc.images, storage="f16")
    b = pm.batch(jobs, p)
    b.run()
    pm.synchronize()
    b.close()
    errs, fills = [], []
    maps = {}
    for r, _ in jobs:
        d = pm.depth_map(r)
        maps[r] = d
        t = sc.depths[r]
        m = np.isfinite(d["depth"]) & (t > 0)
        fills.append(np.isfinite(d["depth"][t > 0]).mean())
        errs.append(np.abs(d["depth"][m] - t[m]) / t[m])
    eb = float(np.median(np.concatenate(errs)))
    for r, srcs in jobs[:2]:
        pm.patchmatch(r, srcs, p)
        s = pm.depth_map(r)
        for k in ("depth", "normals", "cost", "support"):
            assert np.array_equal(s[k], maps[r][k], equal_nan=True), (r, k)
    # the plain FP32 path on the same jobs
    pf = plane.PlaneMVS(sc.cams, sc.images, storage="f32")
    errs = []
    for r, srcs in jobs:
        pf.patchmatch(r, srcs, p)
        d = pf.depth_map(r)
        t = sc.depths[r]
        m = np.isfinite(d["depth"]) & (t > 0)
        errs.append(np.abs(d["depth"][m] - t[m]) / t[m])
    ef = float(np.median(np.concatenate(errs)))
    say("median rel error batch f16", eb, "plain f32", ef, "min fill", min(fills))
    assert min(fills) > 0.9
    assert eb <= ef * 1.05 + 1e-6
