"""Synthetic multi-view scenes with analytic depth (inputs for tests and the bench).

Workloads, by the sizes they are quoted on:
  * ``sphere_arc``    1 reference + 4 sources, 320x240, f=500: sphere r=50 mm at
                       0.35 m over a textured plane at 0.44 m, cameras on a +-15 deg
                       arc about the sphere centre, depth range 0.25-0.45 m.
  * ``sphere_planes`` 9 views, 1920x1440, f=2500: sphere r=50 mm at 0.35 m in front
                       of two tilted planes; 8 cameras on a 100 mm ring around the
                       reference, all aimed at the sphere centre.
  * ``orbit_batch``   N references of uneven resolution (640x480 .. 1920x1440) from
                       a 64-camera orbit around a sphere + box.
  * ``reference_plane`` the reference's own unit-test scene (tests/unit/test_mvs.cpp:
                       112-204), uint8, bit-identical to what that test renders.

Texture: band-limited 3D gradient noise at the world hit point (view-consistent),
box-filtered over ss x ss sub-samples per pixel.  Values are in [0, 1]; reference
mode consumes q = round(255 v) as uint8.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


@dataclass
class Scene:
    cams: list                      # list[N.Camera], index = view id
    images: list                    # list[np.ndarray float32 HxW] in [0,1]
    depths: list                    # list[np.ndarray float32 HxW] truth axial depth (0 = miss)
    min_depth: float
    max_depth: float
    name: str = ""
    meta: dict = field(default_factory=dict)

    def u8(self, i):
        return quantize_u8(self.images[i])

    @property
    def num_views(self):
        return len(self.cams)


def quantize_u8(img: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(np.asarray(img, dtype=np.float64) * 255.0), 0, 255).astype(np.uint8)


def look_at(center, target):
    """World->camera (R, t) with z toward target, x right, y down
    (the axis convention of tests/unit/test_mvs.cpp:150-166)."""
    center = np.asarray(center, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - center
    f /= np.linalg.norm(f)
    r = np.cross(f, np.array([0.0, -1.0, 0.0]))
    if np.linalg.norm(r) < 1e-9:
        r = np.array([1.0, 0.0, 0.0])
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    R = np.stack([r, d, f])
    t = -R @ center
    return R, t


def _desc(spheres=(), planes=(), boxes=(), freq=100.0, octaves=3, seed=0, contrast=0.6,
          supersample=2, noise_sigma=0.0, noise_seed=0):
    d = N.SceneDesc()
    d.num_spheres, d.num_planes, d.num_boxes = len(spheres), len(planes), len(boxes)
    for i, s in enumerate(spheres):
        for j in range(4):
            d.spheres[i][j] = float(s[j])
    for i, p in enumerate(planes):
        for j in range(6):
            d.planes[i][j] = float(p[j])
    for i, b in enumerate(boxes):
        for j in range(6):
            d.boxes[i][j] = float(b[j])
    d.supersample = int(supersample)
    d.texture_freq = float(freq)
    d.octaves = int(octaves)
    d.seed = int(seed) & 0xFFFFFFFF
    d.contrast = float(contrast)
    d.noise_sigma = float(noise_sigma)
    d.noise_seed = int(noise_seed) & 0xFFFFFFFF
    return d


def render(desc: N.SceneDesc, cam: N.Camera, want_image=True, want_depth=True):
    lib = N.scene_lib()
    h, w = cam.height, cam.width
    img = np.zeros((h, w), np.float32) if want_image else None
    dep = np.zeros((h, w), np.float32) if want_depth else None
    st = lib.recon_scene_render(C.byref(desc), C.byref(cam),
                                N.ptr(img, C.c_float) if img is not None else None,
                                N.ptr(dep, C.c_float) if dep is not None else None)
    if st != N.OK:
        raise RuntimeError(f"recon_scene_render failed: {st}")
    return img, dep


def sphere_arc(width=320, height=240, focal=500.0, arc_deg=15.0, n_src=4, seed=0,
               supersample=3, noise_sigma=0.0, min_depth=0.25, max_depth=0.45):
    """1 reference + n_src sources on an arc about a sphere (r=50 mm at 0.35 m)."""
    sc = np.array([0.0, 0.0, 0.35])
    freq = focal / (0.35 * 12.0)
    views = []
    half = n_src // 2
    angles = [0.0]
    for k in range(n_src):
        j = k - half if k < half else k - half + 1
        angles.append(arc_deg * j / half)
    cams = []
    for a in angles:
        th = math.radians(a)
        c = sc + 0.35 * np.array([math.sin(th), 0.0, -math.cos(th)])
        R, t = look_at(c, sc)
        cams.append(N.Camera.make(focal, focal, width / 2.0, height / 2.0, width, height, R, t))
    imgs, deps = [], []
    for i, cam in enumerate(cams):
        d = _desc(spheres=[(0.0, 0.0, 0.35, 0.05)], planes=[(0.0, 0.0, 0.44, 0.0, 0.0, -1.0)],
                  freq=freq, seed=seed, supersample=supersample, noise_sigma=noise_sigma,
                  noise_seed=(seed * 7919 + i + 1))
        img, dep = render(d, cam)
        imgs.append(img)
        deps.append(dep)
    return Scene(cams, imgs, deps, min_depth, max_depth, name=f"sphere_arc_{width}x{height}",
                 meta=dict(angles=angles, focal=focal, seed=seed, noise_sigma=noise_sigma))


def _planes_desc(freq, seed, supersample):
    na = np.array([0.42, 0.10, -0.90]); na /= np.linalg.norm(na)
    nb = np.array([-0.38, -0.15, -0.90]); nb /= np.linalg.norm(nb)
    return _desc(spheres=[(0.0, 0.0, 0.35, 0.05)],
                 planes=[(0.0, 0.0, 0.47, *na), (0.0, 0.0, 0.47, *nb)],
                 freq=freq, seed=seed, supersample=supersample)


def sphere_planes(width=1920, height=1440, focal=2500.0, n_views=9, baseline=0.10, seed=0,
                  supersample=2, min_depth=0.25, max_depth=0.60):
    """Sphere (r=50 mm at 0.35 m) + two tilted planes; view 0 at the origin and
    n_views-1 views on a `baseline` ring around it, all aimed at the sphere centre."""
    target = np.array([0.0, 0.0, 0.35])
    centers = [np.zeros(3)]
    for k in range(n_views - 1):
        phi = 2.0 * math.pi * k / (n_views - 1)
        centers.append(np.array([baseline * math.cos(phi), baseline * math.sin(phi), 0.0]))
    cams = []
    for c in centers:
        R, t = look_at(c, target)
        cams.append(N.Camera.make(focal, focal, width / 2.0, height / 2.0, width, height, R, t))
    desc = _planes_desc(focal / (0.35 * 12.0), seed, supersample)
    imgs, deps = [], []
    for cam in cams:
        img, dep = render(desc, cam)
        imgs.append(img)
        deps.append(dep)
    return Scene(cams, imgs, deps, min_depth, max_depth, name=f"sphere_planes_{width}x{height}",
                 meta=dict(focal=focal, baseline=baseline, seed=seed))


def orbit_batch(num_refs=48, num_cams=64, seed=0, min_res=(640, 480), max_res=(1920, 1440),
                src_range=(4, 10), supersample=1, radius=0.35):
    """num_refs reference views of uneven 4:3 resolution drawn from a num_cams orbit
    around a sphere + box; each reference gets 4-10 random sources among the
    orbit's neighbours.  Returns (Scene of all cameras, list of (ref, [srcs]))."""
    rng = np.random.default_rng(seed)
    target = np.array([0.0, 0.0, 0.0])
    cams, imgs, deps = [], [], []
    widths = rng.integers(min_res[0] // 16, max_res[0] // 16 + 1, size=num_cams) * 16
    desc = _desc(spheres=[(0.0, 0.0, 0.0, 0.05)],
                 boxes=[(-0.11, -0.04, -0.04, -0.06, 0.04, 0.04), (0.06, -0.03, -0.05, 0.12, 0.05, 0.03)],
                 planes=[(0.0, 0.09, 0.0, 0.0, -1.0, 0.0)],
                 freq=1.0, seed=seed, supersample=supersample)
    for k in range(num_cams):
        az = 2.0 * math.pi * k / num_cams
        el = math.radians(-25.0 + 10.0 * math.sin(3.0 * az))
        c = radius * np.array([math.cos(el) * math.sin(az), math.sin(el), -math.cos(el) * math.cos(az)])
        R, t = look_at(c, target)
        w = int(widths[k])
        h = w * 3 // 4
        f = 1.25 * w
        cam = N.Camera.make(f, f, w / 2.0, h / 2.0, w, h, R, t)
        cams.append(cam)
    # one texture frequency for all views (view-consistent), tied to the mean focal
    desc.texture_freq = float(np.mean([c.fx for c in cams])) / (radius * 12.0)
    for cam in cams:
        img, dep = render(desc, cam)
        imgs.append(img)
        deps.append(dep)
    refs = sorted(rng.choice(num_cams, size=num_refs, replace=False).tolist())
    jobs = []
    for r in refs:
        ns = int(rng.integers(src_range[0], src_range[1] + 1))
        cand = sorted({(r + d) % num_cams for d in range(-6, 7)} - {r})
        srcs = sorted(rng.choice(cand, size=min(ns, len(cand)), replace=False).tolist())
        jobs.append((r, srcs))
    sc = Scene(cams, imgs, deps, radius - 0.15, radius + 0.35, name=f"orbit_batch_{num_refs}",
               meta=dict(seed=seed, num_cams=num_cams))
    return sc, jobs


def perturb_camera(cam: N.Camera, delta) -> N.Camera:
    """A misaligned copy of `cam` for pose-error workloads: its world->camera
    transform moved on the left by the rotation vector delta[:3] (Rodrigues'
    formula) and the translation delta[3:], Y -> Exp(w) Y + tau -- the pose
    perturbation the library's adjoint differentiates."""
    import copy
    w = np.asarray(delta[:3], np.float64)
    th = float(np.linalg.norm(w))
    K = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    if th < 1e-12:
        E = np.eye(3) + K + 0.5 * K @ K
    else:
        E = np.eye(3) + (math.sin(th) / th) * K + ((1.0 - math.cos(th)) / (th * th)) * (K @ K)
    c = copy.copy(cam)
    R2 = E @ cam.R_np()
    t2 = E @ cam.t_np() + np.asarray(delta[3:], np.float64)
    for i in range(9):
        c.R[i] = float(R2.reshape(-1)[i])
    for i in range(3):
        c.t[i] = float(t2[i])
    return c


def reference_plane(num_views=5, baseline_m=0.016):
    """tests/unit/test_mvs.cpp:112-204 restated: 160x120 uint8, f=400, plane at 0.35 m."""
    lib = N.scene_lib()
    cams = (N.Camera * num_views)()
    imgs = np.zeros((num_views, 120, 160), np.uint8)
    truth = np.zeros((120, 160), np.float32)
    st = lib.recon_scene_reference_plane(num_views, baseline_m, cams, N.ptr(imgs, C.c_uint8),
                                         N.ptr(truth, C.c_float))
    if st != N.OK:
        raise RuntimeError("recon_scene_reference_plane failed")
    cams = [cams[i] for i in range(num_views)]
    deps = [truth] + [np.zeros_like(truth) for _ in range(num_views - 1)]
    return Scene(cams, [i.astype(np.float32) / 255.0 for i in imgs], deps, 0.15, 0.80,
                 name="reference_plane", meta=dict(u8=imgs))
