"""CPU tests of the oracle (no GPU): the C restatement against the reference's
own outputs (golden fixtures written by the compiled reference, and the
compiled reference itself where it is present), plus closed forms."""
import numpy as np
import pytest

import oracle
from recon_scanner_amd import _native as N
from recon_scanner_amd import scenes
from tests.helpers import bits_equal, cams_from_array, cfg_from_bytes, load, sort_cloud


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def test_xorshift_stream_reproduces_reference_init():
    """The reference's random init draws one xorshift64* value per pixel in
    row-major order from seed 0x5DEECE66D (patchmatch_stereo.cpp:496, 536-539);
    with iterations = 0 its output depth is float(1/inv) of those draws."""
    g = load("patchmatch_48x32.npz")
    cfg = cfg_from_bytes(g["pm_i0_t1_p7_cfg"])
    depth = g["pm_i0_t1_p7_depth"]
    P = depth.size
    u = oracle.xorshift_stream(0x5DEECE66D, P).astype(np.float64) / 4294967296.0
    inv_min, inv_max = 1.0 / cfg.max_depth_m, 1.0 / cfg.min_depth_m
    inv = inv_min + u * (inv_max - inv_min)
    mine = (1.0 / inv).astype(np.float32).reshape(depth.shape)
    fin = np.isfinite(depth)
    assert fin.sum() > 100
    assert bits_equal(mine[fin], depth[fin])


def test_xorshift_seed_zero_substitution():
    # Random(0) starts from 0x9E3779B97F4A7C15 (patchmatch_stereo.cpp:106)
    assert bits_equal(oracle.xorshift_stream(0, 8), oracle.xorshift_stream(0x9E3779B97F4A7C15, 8))


def test_sample_bilinear_closed_forms():
    lib = oracle.restated()
    import ctypes as C
    w, h = 17, 13
    yy, xx = np.mgrid[0:h, 0:w]
    img = (3 * xx + 5 * yy + 7).astype(np.uint8)      # affine ramp, max 3*16+5*12+7 = 115
    out = C.c_double()
    rng = np.random.default_rng(0)
    for u, v in rng.uniform(0, [w - 1, h - 1], size=(200, 2)):
        assert lib.orc_sample_bilinear(img.ctypes.data, w, h, u, v, C.byref(out)) == 1
        assert abs(out.value - (3 * u + 5 * v + 7)) < 1e-9
    # integer coordinates return the pixel itself (pixel centres at integers)
    assert lib.orc_sample_bilinear(img.ctypes.data, w, h, 4.0, 6.0, C.byref(out)) == 1
    assert out.value == img[6, 4]
    # out of range is rejected, not clamped; the last pixel is in range
    for u, v in ((-1e-9, 3), (3, -1e-9), (w - 1 + 1e-9, 3), (3, h - 1 + 1e-9)):
        assert lib.orc_sample_bilinear(img.ctypes.data, w, h, u, v, C.byref(out)) == 0
    assert lib.orc_sample_bilinear(img.ctypes.data, w, h, w - 1.0, h - 1.0, C.byref(out)) == 1
    assert out.value == img[h - 1, w - 1]


def test_compute_ncc_matches_reference_golden():
    g = load("ncc.npz")
    for patch in (7, 11):
        r = oracle.compute_ncc("orc", g["img_a"], g["img_b"], g["coords"], patch)
        assert bits_equal(np.array([x.valid for x in r], np.int32), g[f"ncc_valid_{patch}"])
        assert bits_equal(np.array([x.ncc for x in r]), g[f"ncc_value_{patch}"])
        assert bits_equal(np.array([x.weak_texture for x in r], np.int32), g[f"ncc_weak_{patch}"])
        assert bits_equal(np.array([x.valid_samples for x in r], np.int32), g[f"ncc_samples_{patch}"])
    r = oracle.compute_ncc("orc", g["img_flat"], g["img_a"], g["coords"][:40], 11)
    assert bits_equal(np.array([x.weak_texture for x in r], np.int32), g["ncc_flat_weak"])


def test_compute_ncc_invariants():
    """test_mvs.cpp:259-293: identical windows -> 1, brightness shift -> 1, flat -> weak."""
    sc = scenes.sphere_arc(width=64, height=64, focal=100.0, n_src=2)
    a = sc.u8(0)
    b = np.clip(a.astype(np.int32) + 40, 0, 255).astype(np.uint8)
    assert (a.astype(np.int32) + 40).max() <= 255 or True
    r = oracle.compute_ncc("orc", a, a, np.array([[32.0, 32.0, 32.0, 32.0]]), 11)[0]
    assert r.valid and abs(r.ncc - 1.0) < 1e-6 and not r.weak_texture
    unsat = (a.astype(np.int32) + 40 <= 255).all()
    if unsat:
        r = oracle.compute_ncc("orc", a, b, np.array([[32.0, 32.0, 32.0, 32.0]]), 11)[0]
        assert r.valid and abs(r.ncc - 1.0) < 1e-6
    flat = np.full_like(a, 128)
    r = oracle.compute_ncc("orc", flat, a, np.array([[32.0, 32.0, 32.0, 32.0]]), 11)[0]
    assert r.weak_texture and not r.valid
    # NCC(a, alpha*a + beta) = -1 for alpha < 0
    neg = (255 - a).astype(np.uint8)
    r = oracle.compute_ncc("orc", a, neg, np.array([[32.0, 32.0, 32.0, 32.0]]), 11)[0]
    assert r.valid and abs(r.ncc + 1.0) < 1e-9


@pytest.mark.parametrize("key", ["pm_i0_t1_p7", "pm_i1_t1_p7", "pm_i2_t3_p5", "pm_i2_t1_p7"])
def test_patchmatch_matches_reference_golden(key):
    g = load("patchmatch_48x32.npz")
    cams = cams_from_array(g["cams"])
    u8 = g["images"]
    cfg = cfg_from_bytes(g[key + "_cfg"])
    f, d, s, c = oracle.orc_patchmatch(u8[0], list(u8[1:]), cams[0], cams[1:], cfg)
    assert f == int(g[key + "_filled"])
    assert bits_equal(d, g[key + "_depth"])
    assert bits_equal(s, g[key + "_support"])
    assert bits_equal(c, g[key + "_conf"])


def test_patchmatch_seeded_matches_reference_golden():
    g = load("patchmatch_48x32.npz")
    cams = cams_from_array(g["cams"])
    u8 = g["images"]
    cfg = cfg_from_bytes(g["seeded_cfg"])
    f, d, s, c = oracle.orc_patchmatch(u8[0], list(u8[1:]), cams[0], cams[1:], cfg,
                                       seeds=g["seeded_points"])
    assert f == int(g["seeded_filled"])
    assert bits_equal(d, g["seeded_depth"]) and bits_equal(s, g["seeded_support"])


def test_refine_matches_reference_golden():
    g = load("patchmatch_48x32.npz")
    cams = cams_from_array(g["cams"])
    u8 = g["images"]
    cfg = cfg_from_bytes(g["refine_cfg"])
    n, d, c = oracle.orc_refine(u8[0], list(u8[1:]), cams[0], cams[1:], cfg,
                                g["pm_i2_t1_p7_depth"], g["pm_i2_t1_p7_conf"])
    assert n == int(g["refine_count"])
    assert bits_equal(d, g["refine_depth"]) and bits_equal(c, g["refine_conf"])


def test_fuse_matches_reference_golden():
    g = load("patchmatch_48x32.npz")
    cams = cams_from_array(g["cams"])
    cfg = cfg_from_bytes(g["fuse_cfg"])
    maps = [(g["refine_depth"], g["pm_i2_t1_p7_support"], g["refine_conf"]),
            (g["pm_i2_t1_p7_depth"], g["pm_i2_t1_p7_support"], g["pm_i2_t1_p7_conf"])]
    out = oracle.fuse("orc", maps, [cams[0], cams[0]], cfg)
    a = sort_cloud(out["positions"], out["normals"], out["confidences"])
    b = sort_cloud(g["fuse_positions"], g["fuse_normals"], g["fuse_confidences"])
    assert len(a[0]) == len(b[0]) > 100
    for x, y in zip(a, b):
        assert bits_equal(x, y)
    # keys come out strictly ascending and hash the voxel of each mean position's members
    assert np.all(np.diff(out["keys"]) > 0)


def test_select_views_matches_reference_golden():
    g = load("select_views.npz")
    cams = cams_from_array(g["cams"])
    cfg = N.config_default()
    cfg.num_views = 3
    assert oracle.select_views("orc", cams[0], cams, cfg) == g["selected"].tolist()


@pytest.mark.skipif(not oracle.reference_available(), reason="reference not compiled here")
@pytest.mark.parametrize("threads", [1, 4])
def test_patchmatch_and_refine_match_live_reference(threads):
    sc = scenes.sphere_arc(width=80, height=64, focal=125.0, n_src=4, seed=3)
    u8 = [sc.u8(i) for i in range(5)]
    cfg = N.config_default()
    cfg.patch_size, cfg.iterations, cfg.num_threads = 5, 2, threads
    cfg.min_depth_m, cfg.max_depth_m = 0.25, 0.45
    r = oracle.ref_patchmatch(u8[0], u8[1:], sc.cams[0], sc.cams, sc.cams, [1, 2, 3, 4], cfg)
    o = oracle.orc_patchmatch(u8[0], u8[1:], sc.cams[0], sc.cams[1:], cfg)
    assert r[0] == o[0]
    for x, y in zip(r[1:], o[1:]):
        assert bits_equal(x, y)
    rr = oracle.ref_refine(u8[0], u8[1:], sc.cams[0], sc.cams, sc.cams, [1, 2, 3, 4], cfg, *r[1:])
    oo = oracle.orc_refine(u8[0], u8[1:], sc.cams[0], sc.cams[1:], cfg, o[1], o[3])
    assert rr[0] == oo[0] and bits_equal(rr[1], oo[1]) and bits_equal(rr[2], oo[2])


@pytest.mark.skipif(not oracle.reference_available(), reason="reference not compiled here")
def test_config_defaults_match_reference():
    ref = N.Config()
    oracle.reference().ref_config_default(ref)
    assert bytes(ref) == bytes(N.config_default())


@pytest.mark.skipif(not oracle.reference_available(), reason="reference not compiled here")
def test_camera_math_with_distortion_matches_reference():
    """PixelToRay / ProjectPoint with Brown-Conrady distortion (test_mvs.cpp:210-253)."""
    import ctypes as C
    cam = N.Camera.make(1250.0, 1250.0, 960.0, 540.0, 1920, 1080,
                        dist=(-0.22, 0.06, -0.005, 0.0008, -0.0006))
    uv = np.stack([40.0 + 9.0 * np.arange(120), 30.0 + 5.0 * np.arange(120)], 1).copy()
    a = np.zeros((120, 3)); b = np.zeros((120, 3))
    oracle.reference().ref_pixel_to_ray(C.byref(cam), 120, uv.ctypes.data, a.ctypes.data)
    oracle.restated().orc_pixel_to_ray_batch(C.byref(cam), 120, uv.ctypes.data, b.ctypes.data)
    assert bits_equal(a, b)
    pu = np.zeros((120, 2)); qu = np.zeros((120, 2))
    ok1 = np.zeros(120, np.int32); ok2 = np.zeros(120, np.int32)
    oracle.reference().ref_project_point(C.byref(cam), 120, a.ctypes.data, pu.ctypes.data, ok1.ctypes.data)
    oracle.restated().orc_project_batch(C.byref(cam), 120, b.ctypes.data, qu.ctypes.data, ok2.ctypes.data)
    assert bits_equal(pu, qu) and bits_equal(ok1, ok2)
    assert np.max(np.hypot(*(pu - uv).T)) < 1e-6   # round trip, test_mvs.cpp:247


def test_bgra_to_gray_restatement_closed_forms():
    """BgraToGray (orchestrator.cpp:53-67): the BT.601 integer weights sum to
    256, so equal channels give the channel back; pure primaries give the
    weights scaled by 255/256; alpha is ignored; B and R are not swapped."""
    v = np.arange(256, dtype=np.uint8)
    eq = np.stack([v, v, v, np.full(256, 7, np.uint8)], -1)[None]
    assert np.array_equal(oracle.bgra_to_gray(eq)[0], v)
    px = np.array([[[255, 0, 0, 0], [0, 255, 0, 0], [0, 0, 255, 0], [0, 0, 0, 255]]], np.uint8)
    assert oracle.bgra_to_gray(px)[0].tolist() == [29 * 255 >> 8, 150 * 255 >> 8, 77 * 255 >> 8, 0]
    rng = np.random.default_rng(0)
    a = rng.integers(0, 256, (37, 53, 4), dtype=np.uint8)
    ref = [[(77 * int(p[2]) + 150 * int(p[1]) + 29 * int(p[0])) // 256 for p in row] for row in a]
    assert np.array_equal(oracle.bgra_to_gray(a), np.array(ref, np.uint8))


def _seed_points(w, h, rng):
    """Seeds over a w x h image: a few per cell in most cells, one cell left
    empty, points on cell borders and the image edge, and points the
    reference drops (outside the image, depth <= 0)."""
    pts = []
    for cy in range(0, h, 32):
        for cx in range(0, w, 32):
            if (cx, cy) == (32, 32):
                continue  # an empty cell: random init there
            for _ in range(int(rng.integers(1, 4))):
                pts.append((cx + rng.uniform(0, min(32, w - cx)), cy + rng.uniform(0, min(32, h - cy)),
                            rng.uniform(0.28, 0.42)))
    pts += [(31.999, 5.0, 0.30), (32.0, 5.0, 0.41), (w - 1e-9, h - 1e-9, 0.33), (-0.5, 3.0, 0.3),
            (w, 3.0, 0.3), (4.0, h + 1.0, 0.3), (10.0, 10.0, 0.0), (12.0, 12.0, -1.0)]
    return np.array(pts, np.float64)


@pytest.mark.skipif(not oracle.reference_available(), reason="reference not compiled here")
def test_seed_cells_match_reference_initialisation():
    """The oracle's restatement of the sparse-point seeding (patchmatch_stereo.cpp:
    499-533) against the reference itself: with 0 iterations the reference
    outputs its initial depth 1 / (1 / cell mean) wherever the initial cost is
    finite, so every seeded pixel's output pins the cell assignment, the drop
    rules and the mean."""
    w, h = 100, 70
    sc = scenes.sphere_arc(width=w, height=h, focal=150.0, n_src=4, seed=5)
    u8 = [sc.u8(i) for i in range(5)]
    cfg = N.config_default()
    cfg.patch_size, cfg.iterations, cfg.num_threads = 5, 0, 1
    cfg.min_depth_m, cfg.max_depth_m = 0.25, 0.45
    seeds = _seed_points(w, h, np.random.default_rng(11))
    _, depth, _, _ = oracle.ref_patchmatch(u8[0], u8[1:], sc.cams[0], sc.cams, sc.cams, [1, 2, 3, 4],
                                           cfg, seeds=seeds)
    cells = oracle.seed_cells(w, h, seeds)
    assert cells.shape == (3, 4) and np.isnan(cells[1, 1]) and np.isfinite(cells).sum() == 11
    ys, xs = np.mgrid[0:h, 0:w]
    cd = cells[np.minimum(ys // 32, 2), np.minimum(xs // 32, 3)]
    seeded = np.isfinite(cd) & np.isfinite(depth)
    want = (1.0 / (1.0 / cd[seeded])).astype(np.float32)
    assert np.array_equal(depth[seeded], want)
    # every seeded cell is covered by the comparison
    for cy in range(3):
        for cx in range(4):
            if np.isfinite(cells[cy, cx]):
                block = seeded[cy * 32:(cy + 1) * 32, cx * 32:(cx + 1) * 32]
                assert block.sum() > block.size // 2
    # the empty cell holds the random init, not a constant
    empty = depth[32:64, 32:64]
    assert np.unique(empty[np.isfinite(empty)]).size > 100
    # the plane-mode init rows: fronto-parallel at the float32 cell depth, NaN elsewhere
    planes = oracle.seeded_init_planes(w, h, seeds)
    assert np.array_equal(planes[..., 0][np.isfinite(cd)], cd[np.isfinite(cd)].astype(np.float32))
    assert np.all(planes[np.isfinite(cd)][:, 1:] == [0.0, 0.0, -1.0])
    assert np.all(np.isnan(planes[~np.isfinite(cd)]))
