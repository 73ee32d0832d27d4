"""CPU tests of the FP64 adjoint of the plane cost (the oracle the GPU adjoint
is checked against) and of the torch.autograd harness around it:

* the oracle's point form (ray-plane intersection, then projection) gives the
  same costs as its homography form, which the reduction tests pin to the
  reference's EvaluateCandidate;
* its (depth, normal) gradients agree with central differences (h = 1e-4
  relative), its pose gradients with central differences of 1e-6 rad / m;
* torch.autograd.gradcheck in float64 passes on the harness (this also
  checks the conversion from the kernels' left-perturbation pose gradient to
  the gradient in pose_delta's coordinates);
* a gradient step on a misaligned source pose lowers the cost.
"""
import numpy as np
import pytest
import torch

import oracle
from recon_scanner_amd import adjoint as ADJ
from recon_scanner_amd import scenes

from helpers import OracleBackend, sphere_normals

SRCS = [1, 2, 3, 4]


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def params(**kw):
    p = oracle.PlaneParams()
    p.patch_size, p.window_stride, p.top_k, p.iterations = 11, 2, 2, 2
    p.seed, p.propagation_reach, p.refine_iterations = 0, 4, 4
    p.min_depth_m, p.max_depth_m = 0.25, 0.45
    p.depth_perturb, p.normal_perturb = 0.1, 0.3
    p.photo_threshold, p.value_scale = 1.0, 1.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.fixture(scope="module")
def scene():
    return scenes.sphere_arc(width=128, height=96, focal=200.0)


def pick_pixels(sc, n, seed, jitter=True):
    """n pixels on the surface away from the border, their true planes, and
    (jitter) planes moved off the truth so that gradients are not ~0."""
    rng = np.random.default_rng(seed)
    nrm = sphere_normals(sc)
    ys, xs = np.nonzero(sc.depths[0] > 0)
    keep = (xs > 8) & (xs < 120) & (ys > 8) & (ys < 88)
    xs, ys = xs[keep], ys[keep]
    pick = rng.choice(len(xs), n, replace=False)
    xs, ys = xs[pick], ys[pick]
    planes = np.concatenate([sc.depths[0][ys, xs][:, None], nrm[ys, xs]], 1).astype(np.float64)
    if jitter:
        planes[:, 0] *= 1.0 + rng.normal(0.0, 0.003, n)
        planes[:, 1:] += rng.normal(0.0, 0.05, (n, 3))
        planes[:, 1:] /= np.linalg.norm(planes[:, 1:], axis=1, keepdims=True)
    return np.stack([xs, ys], 1), planes


@pytest.mark.parametrize("top_k", [2, 0])
def test_point_form_equals_homography_form(scene, top_k):
    xy, planes = pick_pixels(scene, 300, 1)
    p = params(top_k=top_k)
    hyps = np.concatenate([xy, planes], 1).astype(np.float32)
    c_h, nv_h, _ = oracle.plane_eval(scene.images, scene.cams, 0, SRCS, p, hyps)
    pf = hyps[:, 2:].astype(np.float64)
    c_p, nv_p, _, _ = oracle.plane_cost_grad(scene.images, scene.cams, 0, SRCS, p, xy, pf, grad=False)
    assert np.array_equal(c_h < 1e29, c_p < 1e29)
    assert np.array_equal(nv_h, nv_p)
    fin = c_h < 1e29
    assert fin.sum() > 250
    assert np.max(np.abs(c_h[fin] - c_p[fin])) < 1e-12


def test_plane_gradient_matches_central_differences(scene):
    xy, planes = pick_pixels(scene, 80, 2)
    p = params()
    c, _, g, _ = oracle.plane_cost_grad(scene.images, scene.cams, 0, SRCS, p, xy, planes)
    ok = np.nonzero(c < 1e29)[0]
    assert len(ok) > 60
    for k in range(4):
        rel = []
        for q in ok:
            h = 1e-4 * (abs(planes[q, 0]) if k == 0 else 1.0)
            pp, pm = planes[q:q + 1].copy(), planes[q:q + 1].copy()
            pp[0, k] += h
            pm[0, k] -= h
            cp = oracle.plane_cost_grad(scene.images, scene.cams, 0, SRCS, p, xy[q:q + 1], pp, grad=False)[0][0]
            cm = oracle.plane_cost_grad(scene.images, scene.cams, 0, SRCS, p, xy[q:q + 1], pm, grad=False)[0][0]
            fd = (cp - cm) / (2 * h)
            rel.append(abs(fd - g[q, k]) / (abs(fd) + abs(g[q, k]) + 1e-12))
        rel = np.array(rel)
        # samples crossing a pixel boundary inside +-h make a few pixels differ
        assert np.median(rel) < 1e-3, (k, np.median(rel))
        assert np.mean(rel < 2e-2) > 0.85, (k, np.sort(rel)[-10:])


def test_normal_gradient_is_tangent(scene):
    """The cost depends on the normal's direction only: dC/dn . n = 0."""
    xy, planes = pick_pixels(scene, 100, 3)
    c, _, g, _ = oracle.plane_cost_grad(scene.images, scene.cams, 0, SRCS, params(), xy, planes)
    ok = c < 1e29
    radial = np.abs(np.sum(g[ok, 1:] * planes[ok, 1:], 1))
    assert np.all(radial <= 1e-9 * (1.0 + np.linalg.norm(g[ok, 1:], axis=1)))


def test_pose_gradient_matches_central_differences(scene):
    """At poses 0.2 deg / 0.5 mm off (away from the minimum, where kinks of
    the bilinear interpolation would dominate the differences), in
    pose_delta's coordinates."""
    xy, planes = pick_pixels(scene, 300, 4, jitter=False)
    p = params()
    rng = np.random.default_rng(4)
    w = rng.uniform(0.5, 1.5, len(xy))
    be = OracleBackend(scene, 0, SRCS, p, xy)
    ax = rng.normal(size=(len(SRCS), 3))
    tr = rng.normal(size=(len(SRCS), 3))
    delta = np.concatenate([np.radians(0.2) * ax / np.linalg.norm(ax, axis=1, keepdims=True),
                            5e-4 * tr / np.linalg.norm(tr, axis=1, keepdims=True)], 1)
    c, _, gs = be.evaluate(planes, delta, w, grad=True)
    gs = ADJ.delta_gradient(gs, delta)
    # Bilinear interpolation is only piecewise smooth: a fraction ~ 2 h f / Z
    # of the samples crosses a pixel boundary within +-h and bends the
    # difference quotient.  Steps of 1e-6 rad / m move samples by ~6e-4 px here.
    h = 1e-6
    for v in range(len(SRCS)):
        fd = np.zeros(6)
        for k in range(6):
            e = np.zeros_like(delta)
            e[v, k] = h
            cp = be.evaluate(planes, delta + e)[0]
            cm = be.evaluate(planes, delta - e)[0]
            m = (cp < 1e29) & (cm < 1e29)
            fd[k] = (np.sum(w[m] * cp[m]) - np.sum(w[m] * cm[m])) / (2 * h)
        assert np.linalg.norm(fd - gs[v]) <= 1e-2 * np.linalg.norm(gs[v]), (v, fd, gs[v])


def test_gradcheck_float64_through_harness(scene):
    xy, planes = pick_pixels(scene, 4, 5)
    be = OracleBackend(scene, 0, SRCS[:2], params(), xy)
    pl = torch.tensor(planes, dtype=torch.float64, requires_grad=True)
    rng = np.random.default_rng(5)
    d0 = np.concatenate([rng.normal(0, 2e-3, (2, 3)), rng.normal(0, 3e-4, (2, 3))], 1)
    pd = torch.tensor(d0, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: ADJ.plane_cost(a, b, be)[0], (pl, pd), eps=1e-6,
                                    atol=1e-5, rtol=1e-3)


def test_left_jacobian_first_order():
    rng = np.random.default_rng(6)
    for _ in range(20):
        phi = rng.normal(0, 0.3, 3)
        h = rng.normal(0, 1e-6, 3)
        lhs = oracle.so3_exp(phi + h)
        rhs = oracle.so3_exp(ADJ.so3_left_jacobian(phi) @ h) @ oracle.so3_exp(phi)
        assert np.max(np.abs(lhs - rhs)) < 1e-11


def test_gradient_step_on_misaligned_pose_lowers_cost(scene):
    """Source 2 off by 0.5 deg / 1 mm: moving its pose against the gradient
    lowers the summed cost of the true planes, and the gradient leans towards
    undoing the misalignment."""
    xy, planes = pick_pixels(scene, 300, 7, jitter=False)
    be = OracleBackend(scene, 0, SRCS, params(), xy)
    rng = np.random.default_rng(7)
    ax = rng.normal(size=3)
    tr = rng.normal(size=3)
    delta = np.zeros((len(SRCS), 6))
    delta[1, :3] = np.radians(0.5) * ax / np.linalg.norm(ax)
    delta[1, 3:] = 1e-3 * tr / np.linalg.norm(tr)
    c1, _, gs = be.evaluate(planes, delta, grad=True)
    g = ADJ.delta_gradient(gs, delta)
    assert g[1] @ delta[1] > 0.0
    c2 = be.evaluate(planes, delta - 1e-6 * g)[0]
    m = (c1 < 1e29) & (c2 < 1e29)
    assert c2[m].sum() < c1[m].sum()
