"""Mixed-precision image storage (FP16 / BF16, FP32 sums) against the FP64
oracle on the unrounded images: the 320 x 240 arc scene with seeded Gaussian
noise sigma = 0.01 on every view, 7 x 7 window, all views.

Rounding a value v in [0, 1] to FP16 (11 significant bits) moves it by at
most 2^-12 v, to BF16 (8 bits) by 2^-9 v.  Spread over a 49-sample window with
texture contrast ~0.1 that bounds the per-view NCC change by ~1e-3 (FP16) and
~1e-2 (BF16) in the worst case; the typical change is far smaller.  Stated
bounds:
  * per-view NCC vs the FP64 oracle on the unrounded images: 99th percentile
    of |dNCC| <= 1e-3 (FP16), 8e-3 (BF16); maximum <= 4e-3 (FP16), 3e-2 (BF16);
  * end to end (3 red-black iterations): the median relative depth error
    against the analytic truth is within 0.1 percentage points of the FP32
    run's.
"""
import numpy as np
import pytest

import oracle
from recon_scanner_amd import plane, scenes

pytestmark = pytest.mark.gpu

BOUNDS = {"f16": (1e-3, 4e-3), "bf16": (8e-3, 3e-2)}


def say(*a):
    print("[case]", *a, flush=True)


def oparams(p):
    o = oracle.PlaneParams()
    for f, _ in oracle.PlaneParams._fields_:
        setattr(o, f, getattr(p, f))
    return o


@pytest.fixture(scope="module")
def noisy():
    return scenes.sphere_arc(noise_sigma=0.01, seed=3)


def params(**kw):
    return plane.default_params(patch_size=7, window_stride=1, top_k=0, iterations=3,
                                min_depth_m=0.25, max_depth_m=0.45, value_scale=1.0, **kw)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_ncc_error_of_storage_format(noisy, storage):
    sc = noisy
    say("noisy arc 320x240 7x7, NCC vs FP64 on unrounded images,", storage)
    rng = np.random.default_rng(11)
    n = 4000
    h, w = sc.depths[0].shape
    xs, ys = rng.integers(4, w - 4, n), rng.integers(4, h - 4, n)
    t = sc.depths[0][ys, xs]
    d = np.where(t > 0, t, 0.4) * rng.uniform(0.99, 1.01, n)
    hyps = np.stack([xs, ys, d, np.zeros(n), np.zeros(n), -np.ones(n)], 1).astype(np.float32)
    p = params()
    pm = plane.PlaneMVS(sc.cams, sc.images, storage=storage)
    _, _, gv = pm.eval_costs(0, [1, 2, 3, 4], p, hyps)
    _, _, ov = oracle.plane_eval(sc.images, sc.cams, 0, [1, 2, 3, 4], oparams(p), hyps)
    both = np.isfinite(gv) & np.isfinite(ov)
    dn = np.abs(gv[both] - ov[both])
    q99, mx = float(np.quantile(dn, 0.99)), float(dn.max())
    say(storage, "views", int(both.sum()), "p50", float(np.median(dn)), "p99", q99, "max", mx)
    assert both.sum() > 2 * n
    assert q99 <= BOUNDS[storage][0] and mx <= BOUNDS[storage][1]


def test_end_to_end_depth_error_within_0p1_points_of_fp32(noisy):
    sc = noisy
    p = params()
    truth = sc.depths[0]
    med = {}
    for storage in ("f32", "f16", "bf16"):
        pm = plane.PlaneMVS(sc.cams, sc.images, storage=storage)
        pm.patchmatch(0, [1, 2, 3, 4], p)
        d = pm.depth_map(0)["depth"]
        m = np.isfinite(d) & (truth > 0)
        med[storage] = float(np.median(np.abs(d[m] - truth[m]) / truth[m]))
    say("median relative depth error", med)
    for s in ("f16", "bf16"):
        assert med[s] <= med["f32"] + 1e-3
