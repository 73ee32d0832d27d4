"""CPU tests of the FP64 plane-mode oracle: its reduction to the reference's
EvaluateCandidate, analytic-truth behaviour, and the counter RNG."""
import math
import os
import re

import numpy as np
import pytest

import oracle
from recon_scanner_amd import _native as N
from recon_scanner_amd import scenes

from helpers import sphere_normals


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def params(**kw):
    p = oracle.PlaneParams()
    p.patch_size, p.window_stride, p.top_k, p.iterations = 7, 1, 0, 2
    p.seed, p.propagation_reach, p.refine_iterations = 0, 4, 4
    p.min_depth_m, p.max_depth_m = 0.25, 0.45
    p.depth_perturb, p.normal_perturb = 0.1, 0.3
    p.photo_threshold, p.value_scale = 1.0, 255.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def lowbias32_py(x, A, B):
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * A) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def test_lowbias32_constants_and_bijectivity():
    lib = oracle.restated()
    A, B = 0x7FEB352D, 0x846CA68B
    hdr = "/root/reference/shaders/common/gpu_algo.h"
    if os.path.exists(hdr):   # the reference's GPU spec names the constants
        txt = open(hdr).read()
        A = int(re.search(r"RECON_GPU_HASH_A\s+(0x[0-9a-fA-F]+)", txt).group(1), 16)
        B = int(re.search(r"RECON_GPU_HASH_B\s+(0x[0-9a-fA-F]+)", txt).group(1), 16)
    rng = np.random.default_rng(0)
    for x in rng.integers(0, 2**32, size=2000, dtype=np.uint64):
        assert lib.orc_lowbias32(int(x)) == lowbias32_py(int(x), A, B)
    # bijection: odd multipliers and xorshifts are invertible -> no collisions
    xs = np.arange(1 << 18, dtype=np.uint32)
    hs = {lib.orc_lowbias32(int(x)) for x in xs[:20000]}
    assert len(hs) == 20000
    # inverse through modular inverses of the multipliers
    Ai, Bi = pow(A, -1, 2**32), pow(B, -1, 2**32)

    def unxorshift(y, s):
        x = y
        for _ in range(32 // s + 1):
            x = y ^ (x >> s)
        return x & 0xFFFFFFFF

    for x in rng.integers(0, 2**32, size=200, dtype=np.uint64):
        h = lib.orc_lowbias32(int(x))
        y = unxorshift(h, 16)
        y = (y * Bi) & 0xFFFFFFFF
        y = unxorshift(y, 15)
        y = (y * Ai) & 0xFFFFFFFF
        y = unxorshift(y, 16)
        assert y == int(x)


def test_rng_uniform_range_and_moments():
    lib = oracle.restated()
    u = np.array([lib.orc_rng_uniform(0, 1, i, k) for i in range(4000) for k in range(4)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005


def translation_scene(w=96, h=72, f=150.0, xs=(-0.012, 0.012, -0.006, 0.006)):
    """Cameras translated along x only (R = I): fronto-parallel planes map by a
    pure translation, the special case where the homography window equals the
    reference's translated window."""
    cams = [N.Camera.make(f, f, w / 2.0, h / 2.0, w, h, np.eye(3), [0, 0, 0])]
    for x in xs:
        cams.append(N.Camera.make(f, f, w / 2.0, h / 2.0, w, h, np.eye(3), [-x, 0.0, 0.0]))
    desc = scenes._desc(spheres=[(0.0, 0.0, 0.35, 0.05)], planes=[(0, 0, 0.44, 0, 0, -1.0)],
                        freq=f / (0.35 * 12.0), seed=5, supersample=2)
    imgs = [scenes.quantize_u8(scenes.render(desc, c)[0]) for c in cams]
    return cams, imgs


def test_reduces_to_reference_evaluate_candidate():
    """n = (0,0,-1), translation-only cameras, k = all views, stride 1 and the
    reference sampling convention: plane cost == EvaluateCandidate cost
    (patchmatch_stereo.cpp:344-418; orc_evaluate is bit-exact with it)."""
    cams, imgs = translation_scene()
    rng = np.random.default_rng(3)
    n = 600
    xs = rng.integers(0, 96, n)
    ys = rng.integers(0, 72, n)
    d = rng.uniform(0.27, 0.44, n).astype(np.float32)
    hyps = np.stack([xs, ys, d, np.zeros(n), np.zeros(n), -np.ones(n)], 1).astype(np.float32)
    for patch in (5, 7):
        p = params(patch_size=patch)
        cost, nv, _ = oracle.plane_eval([i.astype(np.float32) for i in imgs], cams, 0, [1, 2, 3, 4],
                                        p, hyps, sample_shift=0.0)
        cfg = N.config_default()
        cfg.patch_size, cfg.min_depth_m, cfg.max_depth_m = patch, 0.25, 0.45
        uvi = np.stack([xs + 0.5, ys + 0.5, 1.0 / d.astype(np.float64)], 1)
        ref = oracle.orc_evaluate(imgs[0], imgs[1:], cams[0], cams[1:], cfg, uvi)
        rc = np.array([r.cost for r in ref])
        rv = np.array([r.valid_views for r in ref])
        assert np.array_equal(nv, rv)
        fin = rc < 1e29
        assert fin.sum() > 400
        assert np.array_equal(cost < 1e29, fin)
        assert np.max(np.abs(cost[fin] - rc[fin])) < 1e-9


def test_identity_view_gives_perfect_ncc():
    """A source identical to the reference (same pose, same image): H = I for
    every plane, so every valid view has NCC = 1."""
    sc = scenes.sphere_arc(width=64, height=48, focal=100.0)
    cams = [sc.cams[0], sc.cams[0]]
    imgs = [sc.images[0], sc.images[0]]
    rng = np.random.default_rng(1)
    n = 300
    nrm = rng.normal(size=(n, 3))
    nrm[:, 2] = -np.abs(nrm[:, 2]) - 0.5
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    hyps = np.concatenate([rng.integers(6, 58, (n, 1)), rng.integers(6, 42, (n, 1)),
                           rng.uniform(0.26, 0.44, (n, 1)), nrm], 1).astype(np.float32)
    _, nv, vn = oracle.plane_eval(imgs, cams, 0, [1], params(value_scale=1.0, patch_size=11,
                                                             window_stride=2), hyps)
    assert (nv == 1).all()
    assert np.max(np.abs(vn - 1.0)) < 1e-12


def test_true_planes_score_best():
    """Analytic scene: the true plane (depth and normal) costs less than the
    same normal at +-2 % depth; this fixes the plane-induced term b m^T."""
    sc = scenes.sphere_arc(width=128, height=96, focal=200.0)
    nrm = sphere_normals(sc)
    rng = np.random.default_rng(2)
    ys, xs = np.nonzero(sc.depths[0] > 0)
    pick = rng.choice(len(xs), 400, replace=False)
    xs, ys = xs[pick], ys[pick]
    keep = (xs > 8) & (xs < 120) & (ys > 8) & (ys < 88)
    xs, ys = xs[keep], ys[keep]
    d = sc.depths[0][ys, xs]
    p = params(patch_size=11, window_stride=2, top_k=2, value_scale=1.0)
    base = np.concatenate([xs[:, None], ys[:, None], d[:, None], nrm[ys, xs]], 1).astype(np.float32)
    c0, _, _ = oracle.plane_eval(sc.images, sc.cams, 0, [1, 2, 3, 4], p, base)
    worse = 0
    for f in (0.98, 1.02):
        b = base.copy()
        b[:, 2] *= f
        c1, _, _ = oracle.plane_eval(sc.images, sc.cams, 0, [1, 2, 3, 4], p, b)
        worse += np.sum(c1 > c0)
    assert np.median(c0) < 0.05
    assert worse >= 0.9 * 2 * len(xs)   # windows straddling the silhouette excepted


def test_plane_patchmatch_converges_on_sphere_arc():
    sc = scenes.sphere_arc(width=80, height=60, focal=125.0)
    p = params(patch_size=7, window_stride=1, top_k=2, iterations=3, value_scale=1.0)
    planes, cost, sup = oracle.plane_patchmatch(sc.images, sc.cams, 0, [1, 2, 3, 4], p)
    d = planes[..., 0]
    m = (cost < 1e29) & (sc.depths[0] > 0)
    err = np.abs(d[m] - sc.depths[0][m])
    assert m.sum() > 0.9 * d.size
    assert np.median(err) < 0.002


def test_geo_consistency_exact_maps_agree():
    """Exact depth maps are consistent in every view that sees the point."""
    sc = scenes.sphere_planes(width=160, height=120, focal=208.0, n_views=3, supersample=1)
    maps = [np.where(d > 0, d, np.nan).astype(np.float32) for d in sc.depths]
    gs = oracle.geo_consistency(maps, sc.cams, 0.01)
    fin = np.isfinite(maps[0])
    assert (gs[0][~fin] == 0).all()
    assert np.mean(gs[0][fin] >= 2) > 0.9
    # a map scaled by 3 % disagrees with the others at a 1 % threshold
    bad = [maps[0] * 1.03] + maps[1:]
    gb = oracle.geo_consistency(bad, sc.cams, 0.01)
    assert np.mean(gb[0][fin] >= 2) < 0.05


@pytest.mark.parametrize("top_k", [1, 2, 4, 0])
def test_topk_aggregation_brute_force(top_k):
    """The aggregated cost is the mean of the k smallest per-view costs
    1 - NCC over the valid views (docs/05_单目管线.md:1596-1599; k = 0: every
    valid view, the reduction to EvaluateCandidate's mean), rejected above
    3 x photo_threshold.  Recomputed here by sorting the oracle's own per-view
    NCCs; hypotheses near the image border and steep planes leave some views
    invalid, so the k-of-valid rule (fewer valid views than k) is exercised
    too.  A wrong k, divisor, sort order or rejection test fails this."""
    sc = scenes.sphere_planes(width=160, height=120, focal=208.0, n_views=9, supersample=1)
    srcs = list(range(1, 9))
    rng = np.random.default_rng(11)
    n = 1500
    xs = np.concatenate([rng.integers(0, 160, n - 300), rng.integers(0, 6, 150),
                         rng.integers(154, 160, 150)])
    ys = rng.integers(0, 120, n)
    nrm = rng.normal(size=(n, 3))
    nrm[:, 2] = -np.abs(nrm[:, 2]) - 0.05
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.uniform(0.26, 0.58, n)
    t = sc.depths[0][ys, xs]
    half = n // 2
    d[:half] = np.where(t[:half] > 0, t[:half], d[:half])
    nrm[:half] = [0.0, 0.0, -1.0]
    hyps = np.concatenate([xs[:, None], ys[:, None], d[:, None], nrm], 1).astype(np.float32)
    thr = 0.3
    p = params(patch_size=11, window_stride=2, top_k=top_k, value_scale=1.0, min_depth_m=0.25,
               max_depth_m=0.60, photo_threshold=thr)
    cost, nv, vn = oracle.plane_eval(sc.images, sc.cams, 0, srcs, p, hyps)
    k = top_k if top_k else len(srcs)
    nvalid_seen = set()
    rejected = 0
    for q in range(n):
        c = np.sort(1.0 - vn[q][np.isfinite(vn[q])])
        assert nv[q] == len(c)
        nvalid_seen.add(len(c))
        if len(c) == 0:
            assert cost[q] >= 1e29
            continue
        m = min(k, len(c))
        want = sum(float(v) for v in c[:m]) / m
        if want > 3.0 * thr:
            assert cost[q] >= 1e29, (q, want)
            rejected += 1
        else:
            assert abs(cost[q] - want) <= 1e-12, (q, cost[q], want)
    # both regimes of the k-of-valid rule and the rejection occur
    assert min(nvalid_seen) < k <= max(nvalid_seen) or k == len(srcs)
    assert len(nvalid_seen) >= 3
    assert rejected > 0


@pytest.mark.parametrize("delta", [0.75, -1.25, 1.875, -2.0, 0.0])
def test_refine_parabola_vertex_exact(delta):
    """The 5-point least-squares parabola through samples of an exact
    quadratic A (x - m)^2 is that quadratic, so its vertex -- scaled by the
    sample step (inv_per_px / 2) -- lands on m.  Dyadic inv0, step and m make
    every sample position and the vertex exact in FP32: the search returns m
    bit for bit with golden section off.  The reference's index-unit vertex
    (patchmatch_stereo.cpp:898, cand = inv0 + vertex) or a wrong sign of the
    linear term would land elsewhere."""
    inv0, ipp = np.float32(3.0), np.float32(2.0 ** -8)
    m = float(inv0) + delta * float(ipp) / 2.0
    best, cost, n = oracle.refine_1d_closed_form(0, m, 5.0, inv0, ipp, 0, 0.5, 10.0)
    if delta == 0.0:
        # the vertex is the start: no better cost, the start is kept
        assert best == inv0 and cost == 0.0 and n == 5
    else:
        assert best == np.float32(m) and cost == 0.0
        assert n == 5   # 4 fit samples + the vertex
    # a vertex beyond the samples (|x| > 2 steps) is not evaluated, and the
    # fit samples themselves are not candidates: the start is kept
    far = float(inv0) + 2.5 * float(ipp) / 2.0
    best, cost, n = oracle.refine_1d_closed_form(0, far, 5.0, inv0, ipp, 0, 0.5, 10.0)
    assert n == 4 and best == inv0 and cost == 5.0 * (far - float(inv0)) ** 2


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("iters", [1, 2, 4, 8, 12])
def test_refine_golden_section_bracket(kind, iters):
    """Golden section (Kiefer) on a unimodal cost over best +- inv_per_px
    keeps the minimiser inside a bracket that shrinks by 0.618 per step: the
    result is within inv_per_px * 2 * 0.618^N of the minimiser m, with 2 + N
    evaluations after the fit's 4 (+1 for an in-range vertex).  |x - m| (kind
    1) is not smooth, so the parabola cannot find it; the Gaussian dip (kind
    2) is smooth but not quadratic."""
    inv0, ipp = np.float32(2.5), np.float32(0.004)
    rng = np.random.default_rng(kind * 100 + iters)
    for _ in range(40):
        m = float(inv0) + rng.uniform(-0.95, 0.95) * float(ipp)
        best, cost, n = oracle.refine_1d_closed_form(kind, m, 0.01 if kind == 2 else 3.0, inv0,
                                                     ipp, iters, 0.5, 10.0)
        bound = 2.0 * float(ipp) * 0.6180339887 ** iters
        assert abs(float(best) - m) <= bound * 1.001 + 4e-7, (m, best, bound)
        assert n - iters - 2 in (4, 5)
    # a minimiser beyond inv_max: the fit reaches out of range (no vertex),
    # the section is clamped to [inv_min, inv_max] and ends at the edge
    hi = np.float32(float(inv0) + 0.25 * float(ipp))
    best, _, n = oracle.refine_1d_closed_form(kind, float(inv0) + 0.5 * float(ipp),
                                              3.0 if kind == 1 else 0.01, inv0, ipp, iters, 0.5, hi)
    assert n == 4 + 2 + iters
    assert float(hi) - 1.25 * float(ipp) * 0.6180339887 ** iters - 4e-7 <= float(best) <= float(hi)


def test_refine_lowers_costs_on_a_depth_map():
    """On the oracle's own PatchMatch result the refinement never raises a
    pixel's cost, keeps invalid pixels as they are, moves depths by at most
    inv_per_px in inverse depth (the golden-section range) and lowers the
    median depth error against the analytic truth."""
    sc = scenes.sphere_arc(width=80, height=60, focal=125.0)
    srcs = [1, 2, 3, 4]
    p = params(patch_size=7, window_stride=1, top_k=2, iterations=2, value_scale=1.0)
    planes, cost, sup = oracle.plane_patchmatch(sc.images, sc.cams, 0, srcs, p)
    rp, rc, rs = oracle.plane_refine(sc.images, sc.cams, 0, srcs, p, planes, cost, sup)
    valid = cost < 1e29
    assert np.all(rc[valid] <= cost[valid])
    assert np.array_equal(rp[~valid], planes[~valid]) and np.array_equal(rc[~valid], cost[~valid])
    assert np.array_equal(rp[..., 1:], planes[..., 1:])          # normals held
    c = sc.cams
    base = np.mean([np.linalg.norm(c[k].center() - c[0].center()) for k in srcs])
    ipp = 1.0 / (0.5 * (c[0].fx + c[0].fy) * base)
    dinv = np.abs(1.0 / rp[..., 0][valid] - 1.0 / planes[..., 0][valid])
    assert dinv.max() <= 2.0 * ipp * 1.0001
    assert np.mean(rc[valid] < cost[valid]) > 0.5
    truth = sc.depths[0]
    m = valid & (truth > 0)
    e0 = np.median(np.abs(planes[..., 0][m] - truth[m]))
    e1 = np.median(np.abs(rp[..., 0][m] - truth[m]))
    assert e1 < e0
