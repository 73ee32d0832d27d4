"""GPU adjoint of the plane cost (recon_mvs_hip_plane_cost_grad, FP32 sums and
FP64 pose reduction) against the FP64 oracle, central differences of the GPU
cost itself, the torch harness, and the gradient polish.

Tolerances (stated per test):
  * forward cost of the adjoint vs the sweep kernels' cost of the same planes:
    |dc| <= 1e-5 max(1, |c|) at the 99.9th percentile;
  * per-pixel (dD, dn) vs the oracle: relative error (to the oracle vector's
    norm) below 1e-4 at the median and 5e-3 at the 99th percentile (FP32
    sample positions land on the other side of a pixel boundary than FP64
    ones for a few samples);
  * pose gradient (sum over every pixel) vs the oracle: within 1e-3 of the
    per-view vector norm;
  * GPU central differences, depth h = 1e-4 relative: median relative error
    below 2e-2; pose h = 1e-5 rad / m: within 5e-2 of the per-view norm.
"""
import numpy as np
import pytest
import torch

import oracle
from recon_scanner_amd import adjoint as ADJ
from recon_scanner_amd import plane, scenes

pytestmark = pytest.mark.gpu

SRCS = [1, 2, 3, 4, 5, 6]


def say(*a):
    print("[case]", *a, flush=True)


def oparams(p):
    o = oracle.PlaneParams()
    for f, _ in oracle.PlaneParams._fields_:
        setattr(o, f, getattr(p, f))
    return o


def scene_normals(sc, view=0):
    """Normals of the sphere_planes scene in the camera frame of `view`, from
    the depth map's finite differences (exact on the planes, close on the
    sphere)."""
    cam = sc.cams[view]
    d = sc.depths[view].astype(np.float64)
    h, w = d.shape
    yy, xx = np.mgrid[0:h, 0:w]
    P = np.stack([(xx + 0.5 - cam.cx) / cam.fx * d, (yy + 0.5 - cam.cy) / cam.fy * d, d], -1)
    gx = np.zeros_like(P)
    gy = np.zeros_like(P)
    gx[:, 1:-1] = P[:, 2:] - P[:, :-2]
    gy[1:-1] = P[2:] - P[:-2]
    n = np.cross(gy, gx)
    n /= np.maximum(np.linalg.norm(n, axis=-1, keepdims=True), 1e-12)
    n = np.where(n[..., 2:3] > 0, -n, n)
    return n.astype(np.float32)


@pytest.fixture(scope="module")
def setup():
    sc = scenes.sphere_planes(width=480, height=360, focal=625.0, n_views=7, supersample=1)
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, min_depth_m=sc.min_depth,
                             max_depth_m=sc.max_depth, value_scale=1.0, iterations=3)
    rng = np.random.default_rng(0)
    d = sc.depths[0].copy()
    miss = d <= 0
    d[miss] = 0.5
    d = (d * (1.0 + rng.normal(0.0, 0.002, d.shape))).astype(np.float32)
    n = scene_normals(sc)
    n[miss] = [0.0, 0.0, -1.0]
    n = n + rng.normal(0.0, 0.03, n.shape).astype(np.float32)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    pm = plane.PlaneMVS(sc.cams, sc.images, storage="f32")
    pm.set_state(0, SRCS, p, d, n)
    return sc, p, pm, d, n.astype(np.float32)


def test_forward_cost_matches_sweep_cost(setup):
    sc, p, pm, d, n = setup
    say("adjoint forward vs score kernel, 480x360, 6 sources")
    ref = pm.depth_map(0)["cost"]
    c, _, _ = pm.cost_grad(0, SRCS, p, grad=False)
    agree = np.mean((c < 1e29) == (ref < 1e29))
    fin = (c < 1e29) & (ref < 1e29)
    dc = np.abs(c[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin]))
    say("validity agreement", agree, "p99.9 |dc|", float(np.quantile(dc, 0.999)))
    assert agree > 0.999
    assert fin.sum() > 0.8 * c.size
    assert np.quantile(dc, 0.999) <= 1e-5


def test_gradients_match_oracle(setup):
    sc, p, pm, d, n = setup
    say("adjoint vs FP64 oracle, 480x360, 6 sources")
    rng = np.random.default_rng(1)
    w = rng.uniform(0.5, 1.5, d.shape).astype(np.float32)
    c, gp, gs = pm.cost_grad(0, SRCS, p, weights=w)
    h, wd = d.shape
    yy, xx = np.mgrid[0:h, 0:wd]
    xy = np.stack([xx.ravel(), yy.ravel()], 1)
    planes = np.concatenate([d.reshape(-1, 1), n.reshape(-1, 3)], 1).astype(np.float64)
    oc, _, og, os_ = oracle.plane_cost_grad(sc.images, sc.cams, 0, SRCS, oparams(p), xy, planes,
                                            weights=w.reshape(-1).astype(np.float64))
    fin = (c.reshape(-1) < 1e29) & (oc < 1e29)
    g1 = gp.reshape(-1, 4)[fin].astype(np.float64)
    g2 = og[fin]
    # depth and normal parts separately, relative to each oracle vector's norm
    for sl, name in ((slice(0, 1), "depth"), (slice(1, 4), "normal")):
        nrm = np.linalg.norm(g2[:, sl], axis=1)
        big = nrm > 1e-6 * max(1.0, float(np.median(nrm)))
        rel = np.linalg.norm(g1[big, sl] - g2[big, sl], axis=1) / nrm[big]
        say(name, "pixels", int(big.sum()), "median rel", float(np.median(rel)),
            "p99", float(np.quantile(rel, 0.99)))
        assert np.median(rel) < 1e-4
        assert np.quantile(rel, 0.99) < 5e-3
    for v in range(len(SRCS)):
        e = np.linalg.norm(gs[v] - os_[v]) / np.linalg.norm(os_[v])
        say("pose view", v, "rel", float(e))
        assert e < 1e-3


def test_gpu_central_differences(setup):
    sc, p, pm, d, n = setup
    say("GPU central differences, depth and pose")
    c, gp, gs = pm.cost_grad(0, SRCS, p)
    rng = np.random.default_rng(2)
    fin = np.argwhere(c < 1e29)
    pick = fin[rng.choice(len(fin), 400, replace=False)]
    dp, dm = d.copy(), d.copy()
    hstep = 1e-4 * d
    dp[pick[:, 0], pick[:, 1]] += hstep[pick[:, 0], pick[:, 1]]
    dm[pick[:, 0], pick[:, 1]] -= hstep[pick[:, 0], pick[:, 1]]
    # the picked pixels are evaluated together: costs are per pixel
    pm.set_state(0, SRCS, p, dp, n)
    cp = pm.cost_grad(0, SRCS, p, grad=False)[0]
    pm.set_state(0, SRCS, p, dm, n)
    cm = pm.cost_grad(0, SRCS, p, grad=False)[0]
    pm.set_state(0, SRCS, p, d, n)
    r, q = pick[:, 0], pick[:, 1]
    ok = (cp[r, q] < 1e29) & (cm[r, q] < 1e29)
    fd = (cp[r, q] - cm[r, q]).astype(np.float64) / (2.0 * hstep[r, q])
    an = gp[r, q, 0]
    rel = np.abs(fd - an)[ok] / (np.abs(fd) + np.abs(an) + 1e-9)[ok]
    say("depth FD median rel", float(np.median(rel)))
    assert np.median(rel) < 2e-2
    hp = 1e-5
    for v in range(len(SRCS)):
        fdv = np.zeros(6)
        for k in range(6):
            e = np.zeros((len(SRCS), 6))
            e[v, k] = hp
            a = pm.cost_grad(0, SRCS, p, pose_delta=e, grad=False)[0].astype(np.float64)
            b = pm.cost_grad(0, SRCS, p, pose_delta=-e, grad=False)[0].astype(np.float64)
            m = (a < 1e29) & (b < 1e29)
            fdv[k] = (a[m].sum() - b[m].sum()) / (2 * hp)
        err = np.linalg.norm(fdv - gs[v]) / np.linalg.norm(gs[v])
        say("pose view", v, "FD rel", float(err))
        assert err < 5e-2


def test_torch_harness_gpu_backend(setup):
    sc, p, pm, d, n = setup
    say("torch harness on the GPU backend")
    be = ADJ.GpuBackend(pm, 0, SRCS, p)
    planes = torch.tensor(np.concatenate([d[..., None], n], -1), requires_grad=True)
    pose = torch.zeros((len(SRCS), 6), dtype=torch.float64, requires_grad=True)
    cost, valid = ADJ.plane_cost(planes, pose, be)
    loss = (cost * valid).sum()
    loss.backward()
    c, gp, gs = pm.cost_grad(0, SRCS, p)
    assert np.array_equal(valid.numpy(), c < 1e29)
    assert np.array_equal(planes.grad.numpy(), gp)
    assert np.allclose(pose.grad.numpy(), gs, rtol=0, atol=0)


def test_polish_lowers_costs_and_keeps_depth(setup):
    sc, p, pm, d, n = setup
    say("patchmatch + 10-step gradient polish")
    pm.patchmatch(0, SRCS, p)
    before = pm.depth_map(0)
    pm.polish(0, SRCS, p, steps=10)
    after = pm.depth_map(0)
    c0, c1 = before["cost"], after["cost"]
    assert np.all(c1 <= c0)
    improved = np.mean(c1[c0 < 1e29] < c0[c0 < 1e29])
    t = sc.depths[0]
    m = (t > 0) & np.isfinite(before["depth"]) & np.isfinite(after["depth"])
    e0 = np.median(np.abs(before["depth"][m] - t[m]))
    e1 = np.median(np.abs(after["depth"][m] - t[m]))
    say("improved fraction", float(improved), "median |dZ| before", float(e0), "after", float(e1))
    assert improved > 0.2
    assert e1 <= e0 * 1.02
    pm.set_state(0, SRCS, p, d, n)


def test_pose_refinement_recovers_misaligned_source():
    """Source 3 misaligned by 0.5 deg / 1 mm (its images rendered at the true
    pose): gradient descent on its pose_delta through the harness, with the
    true planes held fixed, removes most of the misalignment."""
    say("pose refinement through the torch harness, 480x360")
    sc = scenes.sphere_planes(width=480, height=360, focal=625.0, n_views=7, supersample=1)
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, min_depth_m=sc.min_depth,
                             max_depth_m=sc.max_depth, value_scale=1.0)
    rng = np.random.default_rng(3)
    ax = rng.normal(size=3)
    tr = rng.normal(size=3)
    err = np.concatenate([np.radians(0.5) * ax / np.linalg.norm(ax), 1e-3 * tr / np.linalg.norm(tr)])
    cams = list(sc.cams)
    cams[3] = oracle.perturb_camera(sc.cams[3], err)
    d = sc.depths[0].copy()
    miss = d <= 0
    d[miss] = 0.5
    n = scene_normals(sc)
    n[miss] = [0.0, 0.0, -1.0]
    pm = plane.PlaneMVS(cams, sc.images, storage="f32")
    be = ADJ.GpuBackend(pm, 0, SRCS, p)
    planes = torch.tensor(np.concatenate([d[..., None], n], -1).astype(np.float32))
    # only source 3 (row 2 of pose_delta) moves; translation in units of 4 mm
    rot = torch.zeros(3, dtype=torch.float64, requires_grad=True)
    trn = torch.zeros(3, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([rot, trn], lr=1e-3)
    losses = []

    def pose_of(r, t):
        rows = [torch.zeros(6, dtype=torch.float64) for _ in SRCS]
        rows[2] = torch.cat([r, 4e-3 * t])
        return torch.stack(rows)

    for _ in range(80):
        opt.zero_grad()
        cost, valid = ADJ.plane_cost(planes, pose_of(rot, trn), be)
        loss = (cost * valid).sum() / valid.sum()
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    pose = pose_of(rot, trn)
    # residual misalignment: Exp(pose) applied to the misaligned camera vs truth
    fixed = oracle.perturb_camera(cams[3], pose.detach().numpy()[2])
    R_err = fixed.R_np() @ sc.cams[3].R_np().T
    ang = np.degrees(np.arccos(np.clip((np.trace(R_err) - 1) / 2, -1, 1)))
    c_fix = fixed.center()
    c_true = sc.cams[3].center()
    c_bad = cams[3].center()
    say("loss", losses[0], "->", losses[-1], "angle err deg", float(ang),
        "centre err mm", 1e3 * float(np.linalg.norm(c_fix - c_true)),
        "(was", 1e3 * float(np.linalg.norm(c_bad - c_true)), ")")
    assert losses[-1] < losses[0]
    assert ang < 0.25
