"""GPU parity of reference mode: the HIP library against the reference's own
outputs (golden fixtures) and against the oracle, bit for bit."""
import sys

import numpy as np
import pytest

import oracle
from recon_scanner_amd import _native as N
from recon_scanner_amd import mvs, scenes
from tests.helpers import bits_equal, cams_from_array, cfg_from_bytes, load, sort_cloud

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def pose(c):
    return mvs.CameraPose.from_camera(c)


def intr(c):
    return mvs.CameraIntrinsics.from_camera(c)


def run_pm(u8, cams, cfg, srcs=(1, 2, 3, 4), seeds=None):
    h, w = u8[0].shape
    dm = mvs.DepthMap()
    f = mvs.PatchMatchStereo(u8[0], w, h, [u8[i] for i in srcs], [w] * len(srcs), [h] * len(srcs),
                             pose(cams[0]), [pose(c) for c in cams], intr(cams[0]),
                             [intr(c) for c in cams], list(srcs), cfg, seeds, dm)
    return f, dm


def test_compute_ncc_golden():
    g = load("ncc.npz")
    for patch in (7, 11):
        say("compute_ncc golden patch", patch)
        r = mvs.ComputeNccBatch(g["img_a"], 64, 64, g["img_b"], 64, 64, g["coords"], patch)
        assert bits_equal(np.array([x.valid for x in r], np.int32), g[f"ncc_valid_{patch}"])
        assert bits_equal(np.array([x.ncc for x in r]), g[f"ncc_value_{patch}"])
        assert bits_equal(np.array([x.weak_texture for x in r], np.int32), g[f"ncc_weak_{patch}"])
        assert bits_equal(np.array([x.valid_samples for x in r], np.int32), g[f"ncc_samples_{patch}"])


def test_evaluate_candidates_match_oracle():
    say("evaluate_candidates 96x72 vs oracle")
    sc = scenes.sphere_arc(width=96, height=72, focal=150.0)
    u8 = [sc.u8(i) for i in range(5)]
    cfg = mvs.MvsConfig(patch_size=7, min_depth_m=0.25, max_depth_m=0.45)
    rng = np.random.default_rng(7)
    n = 4000
    q = np.stack([rng.uniform(-2, 98, n), rng.uniform(-2, 74, n), rng.uniform(1.5, 4.5, n)], 1)
    q[:500, 0] = np.floor(q[:500, 0]) + 0.5
    q[:500, 1] = np.floor(q[:500, 1]) + 0.5
    q[500:520, 2] = [0.0, 1e-10, 1e-9, -1.0, 10.0] * 4    # invalid inverse depths
    g = mvs.EvaluateCandidates(u8[0], u8[1:], sc.cams[0], sc.cams[1:], cfg, q)
    o = oracle.orc_evaluate(u8[0], u8[1:], sc.cams[0], sc.cams[1:], cfg, q)
    for f in ("cost", "ncc_sum", "weight_sum"):
        assert bits_equal(np.array([getattr(x, f) for x in g]), np.array([getattr(x, f) for x in o])), f
    for f in ("weak_texture_count", "valid_views"):
        assert [getattr(x, f) for x in g] == [getattr(x, f) for x in o], f
    assert sum(x.cost < 1e29 for x in g) > 1000


@pytest.mark.parametrize("key", ["pm_i0_t1_p7", "pm_i1_t1_p7", "pm_i2_t3_p5", "pm_i2_t1_p7"])
def test_patchmatch_golden(key):
    say("patchmatch golden", key)
    g = load("patchmatch_48x32.npz")
    cams = cams_from_array(g["cams"])
    cfg = cfg_from_bytes(g[key + "_cfg"])
    f, dm = run_pm(list(g["images"]), cams, cfg)
    assert f == int(g[key + "_filled"])
    assert bits_equal(dm.depth.reshape(32, 48), g[key + "_depth"])
    assert bits_equal(dm.support_count.reshape(32, 48), g[key + "_support"])
    assert bits_equal(dm.confidence.reshape(32, 48), g[key + "_conf"])


def test_patchmatch_seeded_golden():
    say("patchmatch seeded golden")
    g = load("patchmatch_48x32.npz")
    cams = cams_from_array(g["cams"])
    cfg = cfg_from_bytes(g["seeded_cfg"])
    f, dm = run_pm(list(g["images"]), cams, cfg, seeds=g["seeded_points"])
    assert f == int(g["seeded_filled"])
    assert bits_equal(dm.depth.reshape(32, 48), g["seeded_depth"])


def test_refine_and_fuse_golden():
    say("refine + fuse golden")
    g = load("patchmatch_48x32.npz")
    cams = cams_from_array(g["cams"])
    u8 = list(g["images"])
    cfg = cfg_from_bytes(g["refine_cfg"])
    dm = mvs.DepthMap()
    dm.Allocate(48, 32)
    dm.depth[:] = g["pm_i2_t1_p7_depth"].reshape(-1)
    dm.support_count[:] = g["pm_i2_t1_p7_support"].reshape(-1)
    dm.confidence[:] = g["pm_i2_t1_p7_conf"].reshape(-1)
    n = mvs.RefineSubpixel(u8[0], 48, 32, u8[1:], [48] * 4, [32] * 4, pose(cams[0]),
                           [pose(c) for c in cams], intr(cams[0]), [intr(c) for c in cams],
                           [1, 2, 3, 4], cfg, dm)
    assert n == int(g["refine_count"])
    assert bits_equal(dm.depth.reshape(32, 48), g["refine_depth"])
    assert bits_equal(dm.confidence.reshape(32, 48), g["refine_conf"])
    cfg = cfg_from_bytes(g["fuse_cfg"])
    dm2 = mvs.DepthMap()
    dm2.Allocate(48, 32)
    dm2.depth[:] = g["pm_i2_t1_p7_depth"].reshape(-1)
    dm2.support_count[:] = g["pm_i2_t1_p7_support"].reshape(-1)
    dm2.confidence[:] = g["pm_i2_t1_p7_conf"].reshape(-1)
    cloud = mvs.FuseDepthMaps([dm, dm2], [pose(cams[0])] * 2, [intr(cams[0])] * 2, cfg)
    a = sort_cloud(cloud.positions, cloud.normals, cloud.confidences)
    b = sort_cloud(g["fuse_positions"], g["fuse_normals"], g["fuse_confidences"])
    assert cloud.size() == len(b[0])
    for x, y in zip(a, b):
        assert bits_equal(x, y)
    assert np.all(np.diff(cloud.voxel_keys) > 0)
    # the same cloud in Poisson input form (UnifiedPointCloud::ToOrientedCloud)
    oc = mvs.FetchOrientedCloud()
    pos, nrm, col, w = oracle.to_oriented_cloud(cloud.positions, cloud.normals, cloud.colors,
                                                cloud.confidences)
    assert oc.size() == cloud.size()
    assert bits_equal(oc.positions, pos) and bits_equal(oc.normals, nrm)
    assert bits_equal(oc.colors, col) and bits_equal(oc.weights, w)


def test_index_contract_and_early_returns():
    say("index contract")
    g = load("patchmatch_48x32.npz")
    cams = cams_from_array(g["cams"])
    u8 = list(g["images"])
    cfg = cfg_from_bytes(g["pm_i1_t1_p7_cfg"])
    dm = mvs.DepthMap()
    # relative intrinsics (size == num sources) give the same result as global
    f_rel = mvs.PatchMatchStereo(u8[0], 48, 32, u8[1:], [48] * 4, [32] * 4, pose(cams[0]),
                                 [pose(c) for c in cams], intr(cams[0]), [intr(c) for c in cams[1:]],
                                 [1, 2, 3, 4], cfg, None, dm)
    assert f_rel == int(g["pm_i1_t1_p7_filled"])
    assert bits_equal(dm.depth.reshape(32, 48), g["pm_i1_t1_p7_depth"])
    # ambiguous intrinsics length, mismatched arrays, bad global index, empty -> 0
    bad = mvs.DepthMap()
    assert mvs.PatchMatchStereo(u8[0], 48, 32, u8[1:], [48] * 4, [32] * 4, pose(cams[0]),
                                [pose(c) for c in cams], intr(cams[0]), [intr(c) for c in cams[:3]],
                                [1, 2, 3, 4], cfg, None, bad) == 0
    assert mvs.PatchMatchStereo(u8[0], 48, 32, u8[1:3], [48] * 4, [32] * 4, pose(cams[0]),
                                [pose(c) for c in cams], intr(cams[0]), [intr(c) for c in cams],
                                [1, 2, 3, 4], cfg, None, bad) == 0
    assert mvs.PatchMatchStereo(u8[0], 48, 32, u8[1:], [48] * 4, [32] * 4, pose(cams[0]),
                                [pose(c) for c in cams], intr(cams[0]), [intr(c) for c in cams],
                                [1, 2, 3, 9], cfg, None, bad) == 0
    assert mvs.PatchMatchStereo(u8[0], 48, 32, [], [], [], pose(cams[0]), [pose(c) for c in cams],
                                intr(cams[0]), [intr(c) for c in cams], [], cfg, None, bad) == 0
    # zero baseline: DepthMap allocated (all NaN), returns 0
    same = [cams[0]] * 5
    z = mvs.DepthMap()
    assert mvs.PatchMatchStereo(u8[0], 48, 32, u8[1:], [48] * 4, [32] * 4, pose(cams[0]),
                                [pose(c) for c in same], intr(cams[0]), [intr(c) for c in same],
                                [1, 2, 3, 4], cfg, None, z) == 0
    assert z.depth.size == 48 * 32 and np.isnan(z.depth).all()


def test_sphere_arc_320x240_matches_oracle_and_truth():
    """cfg: 1 ref + 4 sources, 320x240, 7x7, 3 iterations, num_threads 8."""
    say("sphere_arc 320x240 reference mode vs oracle")
    sc = scenes.sphere_arc()
    u8 = [sc.u8(i) for i in range(5)]
    cfg = mvs.MvsConfig(patch_size=7, iterations=3, num_threads=8, min_depth_m=0.25,
                        max_depth_m=0.45)
    f, dm = run_pm(u8, sc.cams, cfg)
    of, od, os_, oc = oracle.orc_patchmatch(u8[0], u8[1:], sc.cams[0], sc.cams[1:], cfg)
    assert f == of
    assert bits_equal(dm.depth.reshape(240, 320), od)
    assert bits_equal(dm.support_count.reshape(240, 320), os_)
    assert bits_equal(dm.confidence.reshape(240, 320), oc)
    d = dm.depth.reshape(240, 320)
    m = np.isfinite(d) & (sc.depths[0] > 0)
    err = np.abs(d[m] - sc.depths[0][m])
    say("filled", f, "median |dZ| mm", float(np.median(err) * 1000))
    assert f > 0.9 * 320 * 240
    assert np.median(err) < 0.01


def test_reference_unit_test_scene():
    """Replays TestMvsDepthAccuracy (tests/unit/test_mvs.cpp:299-467) on the GPU,
    with its thresholds, and checks the result against the oracle bit for bit."""
    say("reference test_mvs scene")
    sc = scenes.reference_plane(5, 0.016)
    u8 = list(sc.meta["u8"])
    cams = sc.cams
    cfg = mvs.MvsConfig(num_views=4, num_threads=8, patch_size=11, iterations=3,
                        random_search_count=5, random_search_radius_ratio=0.35,
                        subpixel_iterations=5, min_depth_m=0.15, max_depth_m=0.80,
                        fusion_voxel_size_m=0.0005)
    f, dm = run_pm(u8, cams, cfg)
    of, od, os_, oc = oracle.orc_patchmatch(u8[0], u8[1:], cams[0], cams[1:], cfg)
    assert f == of and bits_equal(dm.depth.reshape(120, 160), od)
    assert f > 160 * 120 * 0.5
    fin = np.isfinite(dm.depth)
    assert (fin & (dm.support_count >= 2)).sum() > f * 0.5
    n = mvs.RefineSubpixel(u8[0], 160, 120, u8[1:], [160] * 4, [120] * 4, pose(cams[0]),
                           [pose(c) for c in cams], intr(cams[0]), [intr(c) for c in cams],
                           [1, 2, 3, 4], cfg, dm)
    on, odr, ocr = oracle.orc_refine(u8[0], u8[1:], cams[0], cams[1:], cfg, od, oc)
    assert n == on and bits_equal(dm.depth.reshape(120, 160), odr)
    d = dm.depth.reshape(120, 160)
    truth = sc.depths[0]
    crop = (slice(20, 100), slice(20, 140))
    dd, tt = d[crop], truth[crop]
    m = np.isfinite(dd) & (tt > 0)
    err = np.sort(np.abs(dd[m].astype(np.float64) - tt[m]) * 1000.0)
    median, p95 = err[len(err) // 2], err[int(len(err) * 0.95)]
    say("median mm", median, "p95 mm", p95, "n", len(err))
    assert median <= 19.1 * 0.15 * 1.5 and p95 <= 19.1 * 0.15 * 3.0
    cloud = mvs.FuseDepthMaps([dm], [pose(cams[0])], [intr(cams[0])], cfg)
    assert cloud.size() > 1000


@pytest.mark.parametrize("patch", [9, 4, 10, 6])
def test_ragged_tiles_and_source_sizes(patch):
    """Reference and sources off the pixel kernels' 32x4 tile grid (101x67,
    sources cropped to other sizes: the quads of each source carry its own
    edge clamps), odd and even patches on the generic kernels (9, 4) and on
    the unrolled ones (10 and 6 share 11x11 and 7x7's half-width, with their
    own validity threshold patch^2 / 2), PatchMatchStereo then
    RefineSubpixel, against the oracle bit for bit."""
    say("ragged tiles, patch", patch)
    sc = scenes.sphere_arc(width=101, height=67, focal=160.0)
    u8 = [sc.u8(i) for i in range(5)]
    srcs = [u8[1][:61, :99], u8[2][:67, :90], np.ascontiguousarray(u8[3][:50, :101]), u8[4]]
    srcs = [np.ascontiguousarray(s) for s in srcs]
    sw = [s.shape[1] for s in srcs]
    sh = [s.shape[0] for s in srcs]
    cfg = mvs.MvsConfig(patch_size=patch, iterations=2, num_threads=3, min_depth_m=0.25,
                        max_depth_m=0.45)
    cams = sc.cams
    dm = mvs.DepthMap()
    f = mvs.PatchMatchStereo(u8[0], 101, 67, srcs, sw, sh, pose(cams[0]), [pose(c) for c in cams],
                             intr(cams[0]), [intr(c) for c in cams], [1, 2, 3, 4], cfg, None, dm)
    of, od, os_, oc = oracle.orc_patchmatch(u8[0], srcs, cams[0], cams[1:], cfg)
    say("filled", f, of)
    assert f == of and f > 0
    assert bits_equal(dm.depth.reshape(67, 101), od)
    assert bits_equal(dm.support_count.reshape(67, 101), os_)
    assert bits_equal(dm.confidence.reshape(67, 101), oc)
    n = mvs.RefineSubpixel(u8[0], 101, 67, srcs, sw, sh, pose(cams[0]), [pose(c) for c in cams],
                           intr(cams[0]), [intr(c) for c in cams], [1, 2, 3, 4], cfg, dm)
    on, odr, ocr = oracle.orc_refine(u8[0], srcs, cams[0], cams[1:], cfg, od, oc)
    say("refined", n, on)
    assert n == on and n > 0
    assert bits_equal(dm.depth.reshape(67, 101), odr)
    assert bits_equal(dm.confidence.reshape(67, 101), ocr)


def test_distorted_cameras_match_reference():
    """Cameras with radial and tangential distortion (k1, k2, k3, p1, p2 all
    non-zero, different per view): the device camera arithmetic -- the
    Newton iterations of RemoveDistortion and ApplyDistortion
    (camera.cpp:47-129) inside every projection -- reproduces the reference
    bit for bit: EvaluateCandidate costs against the oracle (itself pinned to
    the compiled reference with distortion on the host), PatchMatchStereo and
    RefineSubpixel outputs against the reference's own code compiled from its
    sources."""
    say("reference mode with distorted cameras, sphere_arc 96x72")
    sc = scenes.sphere_arc(width=96, height=72, focal=150.0)
    u8 = [sc.u8(i) for i in range(5)]
    dists = [(-0.08, 0.02, 0.003, 0.001, -0.0005), (-0.05, 0.01, 0.0, -0.0008, 0.0006),
             (0.04, -0.015, 0.002, 0.0005, 0.0004), (-0.1, 0.03, -0.004, 0.0, 0.001),
             (0.02, 0.0, 0.001, -0.001, -0.0002)]
    cams = [N.Camera.make(c.fx, c.fy, c.cx, c.cy, c.width, c.height, c.R_np(), c.t_np(), d)
            for c, d in zip(sc.cams, dists)]
    cfg = mvs.MvsConfig(patch_size=7, iterations=2, num_threads=3, min_depth_m=0.25,
                        max_depth_m=0.45)
    rng = np.random.default_rng(9)
    q = np.stack([rng.uniform(0, 96, 1500), rng.uniform(0, 72, 1500), rng.uniform(2.3, 3.9, 1500)], 1)
    g = mvs.EvaluateCandidates(u8[0], u8[1:], cams[0], cams[1:], cfg, q)
    o = oracle.orc_evaluate(u8[0], u8[1:], cams[0], cams[1:], cfg, q)
    assert bits_equal(np.array([x.cost for x in g]), np.array([x.cost for x in o]))
    assert sum(x.cost < 1e29 for x in g) > 500
    if not oracle.reference_available():
        pytest.skip("the reference's compiled CPU path (oracle/_ref) is not built")
    srcs = [1, 2, 3, 4]
    f, dm = run_pm(u8, cams, cfg, srcs)
    rf, rd, rs, rc = oracle.ref_patchmatch(u8[0], [u8[i] for i in srcs], cams[0], cams, cams, srcs, cfg)
    say("filled", f, rf)
    assert f == rf and f > 0.5 * 96 * 72
    assert bits_equal(dm.depth.reshape(72, 96), rd)
    assert bits_equal(dm.support_count.reshape(72, 96), rs)
    assert bits_equal(dm.confidence.reshape(72, 96), rc)
    n = mvs.RefineSubpixel(u8[0], 96, 72, [u8[i] for i in srcs], [96] * 4, [72] * 4, pose(cams[0]),
                           [pose(c) for c in cams], intr(cams[0]), [intr(c) for c in cams], srcs, cfg, dm)
    rn, rdr, rcr = oracle.ref_refine(u8[0], [u8[i] for i in srcs], cams[0], cams, cams, srcs, cfg, rd, rs, rc)
    say("refined", n, rn)
    assert n == rn and n > 0
    assert bits_equal(dm.depth.reshape(72, 96), rdr)
    assert bits_equal(dm.confidence.reshape(72, 96), rcr)
