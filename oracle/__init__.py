"""The oracle — TEST INFRASTRUCTURE ONLY.

Python bindings of the two CPU checkers built by oracle/Makefile:

* ``restated()`` -> ``_lib/libpmoracle.so``: my plain-C restatement of the
  reference path (pm_oracle.c) and the FP64 plane-mode restatement
  (pm_oracle_plane.c);
* ``reference()`` -> ``_ref/libreconref.so``: the reference's own CPU MVS code
  (core/src/mvs/patchmatch_stereo.cpp, core/src/util/camera.cpp,
  core/src/util/linalg.cpp) compiled in place, behind ref_shim.cpp.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this package, and only as the checker / the timed CPU baseline; the
product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from recon_scanner_amd import _native as N

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "_lib", "libpmoracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libreconref.so")
REF_SRC = "/root/reference"

_P = C.POINTER
_cam_p = _P(N.Camera)
_cfg_p = _P(N.Config)


def build(with_reference: bool | None = None) -> None:
    """Compile the restatement, and the reference where its sources exist."""
    if with_reference is None:
        with_reference = os.path.isdir(os.path.join(REF_SRC, "core"))
    targets = ["oracle"] + (["ref"] if with_reference else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def _declare_common(lib, prefix):
    f = getattr(lib, prefix + "compute_ncc")
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                  C.c_void_p, C.c_int, _P(N.NccResult)]
    f.restype = None


_restated = None
_reference = None


def restated():
    global _restated
    if _restated is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run oracle.build()")
        lib = C.CDLL(ORACLE_LIB)
        _declare_common(lib, "orc_")
        lib.orc_xorshift_stream.argtypes = [C.c_uint64, C.c_int, C.c_void_p]
        lib.orc_sample_bilinear.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                            _P(C.c_double)]
        lib.orc_sample_bilinear.restype = C.c_int
        lib.orc_evaluate.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _P(C.c_void_p),
                                     C.c_void_p, C.c_void_p, _cam_p, _cam_p, _cfg_p, C.c_int,
                                     C.c_void_p, _P(N.CostResult)]
        lib.orc_patchmatch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _P(C.c_void_p),
                                       C.c_void_p, C.c_void_p, _cam_p, _cam_p, _cfg_p,
                                       C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_patchmatch.restype = C.c_uint32
        lib.orc_refine_subpixel.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _P(C.c_void_p),
                                            C.c_void_p, C.c_void_p, _cam_p, _cam_p, _cfg_p,
                                            C.c_void_p, C.c_void_p]
        lib.orc_refine_subpixel.restype = C.c_uint32
        lib.orc_fuse.argtypes = [C.c_int, C.c_void_p, C.c_void_p, _P(C.c_void_p), _P(C.c_void_p),
                                 _P(C.c_void_p), _cam_p, _cfg_p, C.c_int64, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_fuse.restype = C.c_int64
        lib.orc_select_views.argtypes = [_cam_p, C.c_int, _cam_p, _cfg_p, C.c_void_p, C.c_int]
        lib.orc_select_views.restype = C.c_int
        lib.orc_solve3.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_solve3.restype = C.c_int
        lib.orc_pixel_to_ray_batch.argtypes = [_cam_p, C.c_int, C.c_void_p, C.c_void_p]
        lib.orc_project_batch.argtypes = [_cam_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_lowbias32.argtypes = [C.c_uint32]
        lib.orc_lowbias32.restype = C.c_uint32
        _declare_plane(lib)
        _restated = lib
    return _restated


class PlaneParams(C.Structure):
    """Layout of recon_mvs_plane_params (include/recon_mvs_hip.h)."""
    _fields_ = [("patch_size", C.c_int32), ("window_stride", C.c_int32), ("top_k", C.c_int32),
                ("iterations", C.c_int32), ("seed", C.c_uint32), ("propagation_reach", C.c_int32),
                ("refine_iterations", C.c_int32), ("tile_schedule", C.c_int32),
                ("min_depth_m", C.c_double), ("max_depth_m", C.c_double),
                ("depth_perturb", C.c_double), ("normal_perturb", C.c_double),
                ("photo_threshold", C.c_double), ("value_scale", C.c_double)]


def _declare_plane(lib):
    lib.orc_plane_eval.argtypes = [C.c_int, _P(C.c_void_p), _cam_p, C.c_int, C.c_int, C.c_void_p,
                                   C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]
    lib.orc_plane_eval.restype = C.c_int
    lib.orc_plane_patchmatch.argtypes = [C.c_int, _P(C.c_void_p), _cam_p, C.c_int, C.c_int,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
    lib.orc_plane_patchmatch.restype = C.c_int
    lib.orc_geo_consistency.argtypes = [C.c_int, _P(C.c_void_p), _cam_p, C.c_double, _P(C.c_void_p)]
    lib.orc_geo_consistency.restype = None
    lib.orc_plane_cost_grad.argtypes = [C.c_int, _P(C.c_void_p), _cam_p, C.c_int, C.c_int,
                                        C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.orc_plane_cost_grad.restype = C.c_int
    lib.orc_rng_uniform.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
    lib.orc_rng_uniform.restype = C.c_float
    lib.orc_plane_refine.argtypes = [C.c_int, _P(C.c_void_p), _cam_p, C.c_int, C.c_int, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.orc_plane_refine.restype = C.c_int
    lib.orc_refine_1d_closed_form.argtypes = [C.c_int, C.c_double, C.c_double, C.c_float,
                                              C.c_float, C.c_int, C.c_float, C.c_float,
                                              C.c_void_p, C.c_void_p]
    lib.orc_refine_1d_closed_form.restype = C.c_int


def reference_available() -> bool:
    return os.path.exists(REF_LIB)


def reference():
    global _reference
    if _reference is None:
        if not os.path.exists(REF_LIB):
            raise RuntimeError(f"{REF_LIB} missing: the reference was not compiled here")
        lib = C.CDLL(REF_LIB)
        _declare_common(lib, "ref_")
        lib.ref_config_default.argtypes = [_cfg_p]
        lib.ref_pixel_to_ray.argtypes = [_cam_p, C.c_int, C.c_void_p, C.c_void_p]
        lib.ref_project_point.argtypes = [_cam_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        common = [C.c_void_p, C.c_int, C.c_int, C.c_int, _P(C.c_void_p), C.c_void_p, C.c_void_p,
                  _cam_p, C.c_int, _cam_p, C.c_int, _cam_p, C.c_void_p, _cfg_p]
        lib.ref_patchmatch.argtypes = common + [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                                C.c_void_p]
        lib.ref_patchmatch.restype = C.c_uint32
        lib.ref_refine_subpixel.argtypes = common + [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ref_refine_subpixel.restype = C.c_uint32
        lib.ref_fuse.argtypes = [C.c_int, C.c_void_p, C.c_void_p, _P(C.c_void_p), _P(C.c_void_p),
                                 _P(C.c_void_p), _cam_p, _cfg_p, C.c_int64, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ref_fuse.restype = C.c_int64
        lib.ref_select_views.argtypes = [_cam_p, C.c_int, _cam_p, _cfg_p, C.c_void_p, C.c_int]
        lib.ref_select_views.restype = C.c_int
        _reference = lib
    return _reference


# --------------------------------------------------------------------------
# Convenience wrappers (numpy in, numpy out)
# --------------------------------------------------------------------------

def _ptr_array(arrs):
    return (C.c_void_p * max(len(arrs), 1))(*[a.ctypes.data for a in arrs])


def _cams(cams):
    return (N.Camera * max(len(cams), 1))(*cams)


class _Src:
    def __init__(self, srcs):
        self.srcs = [np.ascontiguousarray(s) for s in srcs]
        self.ptrs = _ptr_array(self.srcs)
        self.w = np.array([s.shape[1] for s in self.srcs], np.int32)
        self.h = np.array([s.shape[0] for s in self.srcs], np.int32)


def compute_ncc(which, a, b, coords, patch):
    lib = reference() if which == "ref" else restated()
    fn = lib.ref_compute_ncc if which == "ref" else lib.orc_compute_ncc
    a = np.ascontiguousarray(a, np.uint8)
    b = np.ascontiguousarray(b, np.uint8)
    coords = np.ascontiguousarray(coords, np.float64)
    n = coords.shape[0]
    out = (N.NccResult * max(n, 1))()
    fn(a.ctypes.data, a.shape[1], a.shape[0], b.ctypes.data, b.shape[1], b.shape[0], n,
       coords.ctypes.data, patch, out)
    return out[:n]


def ref_patchmatch(ref_img, src_imgs, ref_cam, poses, intrinsics, source_indices, cfg,
                   seeds=None):
    """Reference PatchMatchStereo with the full index contract (global poses)."""
    lib = reference()
    s = _Src(src_imgs)
    ref_img = np.ascontiguousarray(ref_img, np.uint8)
    h, w = ref_img.shape
    depth = np.full((h, w), np.nan, np.float32)
    sup = np.zeros((h, w), np.uint8)
    conf = np.zeros((h, w), np.float32)
    idx = np.ascontiguousarray(source_indices, np.int32)
    pcams, icams = _cams(poses), _cams(intrinsics)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.float64)
    filled = lib.ref_patchmatch(ref_img.ctypes.data, w, h, len(s.srcs), s.ptrs, s.w.ctypes.data,
                                s.h.ctypes.data, C.byref(ref_cam), len(poses), pcams,
                                len(intrinsics), icams, idx.ctypes.data, C.byref(cfg),
                                None if sd is None else sd.ctypes.data,
                                0 if sd is None else sd.shape[0], depth.ctypes.data,
                                sup.ctypes.data, conf.ctypes.data)
    return filled, depth, sup, conf


def ref_refine(ref_img, src_imgs, ref_cam, poses, intrinsics, source_indices, cfg, depth, sup,
               conf):
    lib = reference()
    s = _Src(src_imgs)
    ref_img = np.ascontiguousarray(ref_img, np.uint8)
    h, w = ref_img.shape
    depth = np.array(depth, np.float32, copy=True)
    sup = np.array(sup, np.uint8, copy=True)
    conf = np.array(conf, np.float32, copy=True)
    idx = np.ascontiguousarray(source_indices, np.int32)
    pcams, icams = _cams(poses), _cams(intrinsics)
    n = lib.ref_refine_subpixel(ref_img.ctypes.data, w, h, len(s.srcs), s.ptrs, s.w.ctypes.data,
                                s.h.ctypes.data, C.byref(ref_cam), len(poses), pcams,
                                len(intrinsics), icams, idx.ctypes.data, C.byref(cfg),
                                depth.ctypes.data, sup.ctypes.data, conf.ctypes.data)
    return n, depth, conf


def orc_patchmatch(ref_img, src_imgs, ref_cam, src_cams, cfg, seeds=None):
    lib = restated()
    s = _Src(src_imgs)
    ref_img = np.ascontiguousarray(ref_img, np.uint8)
    h, w = ref_img.shape
    depth = np.full((h, w), np.nan, np.float32)
    sup = np.zeros((h, w), np.uint8)
    conf = np.zeros((h, w), np.float32)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.float64)
    filled = lib.orc_patchmatch(ref_img.ctypes.data, w, h, len(s.srcs), s.ptrs, s.w.ctypes.data,
                                s.h.ctypes.data, C.byref(ref_cam), _cams(src_cams), C.byref(cfg),
                                None if sd is None else sd.ctypes.data,
                                0 if sd is None else sd.shape[0], depth.ctypes.data,
                                sup.ctypes.data, conf.ctypes.data)
    return filled, depth, sup, conf


def orc_refine(ref_img, src_imgs, ref_cam, src_cams, cfg, depth, conf):
    lib = restated()
    s = _Src(src_imgs)
    ref_img = np.ascontiguousarray(ref_img, np.uint8)
    h, w = ref_img.shape
    depth = np.array(depth, np.float32, copy=True)
    conf = np.array(conf, np.float32, copy=True)
    n = lib.orc_refine_subpixel(ref_img.ctypes.data, w, h, len(s.srcs), s.ptrs, s.w.ctypes.data,
                                s.h.ctypes.data, C.byref(ref_cam), _cams(src_cams), C.byref(cfg),
                                depth.ctypes.data, conf.ctypes.data)
    return n, depth, conf


def orc_evaluate(ref_img, src_imgs, ref_cam, src_cams, cfg, uvi):
    lib = restated()
    s = _Src(src_imgs)
    ref_img = np.ascontiguousarray(ref_img, np.uint8)
    uvi = np.ascontiguousarray(uvi, np.float64)
    n = uvi.shape[0]
    out = (N.CostResult * max(n, 1))()
    lib.orc_evaluate(ref_img.ctypes.data, ref_img.shape[1], ref_img.shape[0], len(s.srcs), s.ptrs,
                     s.w.ctypes.data, s.h.ctypes.data, C.byref(ref_cam), _cams(src_cams),
                     C.byref(cfg), n, uvi.ctypes.data, out)
    return out[:n]


def _fuse_args(maps, cams):
    ws = np.array([m[0].shape[1] for m in maps], np.int32)
    hs = np.array([m[0].shape[0] for m in maps], np.int32)
    ds = [np.ascontiguousarray(m[0], np.float32) for m in maps]
    ss = [np.ascontiguousarray(m[1], np.uint8) for m in maps]
    cs = [np.ascontiguousarray(m[2], np.float32) for m in maps]
    keep = (ds, ss, cs)
    return ws, hs, _ptr_array(ds), _ptr_array(ss), _ptr_array(cs), _cams(cams), keep


def fuse(which, maps, cams, cfg):
    """maps: list of (depth, support, conf).  Returns dict of numpy arrays."""
    lib = reference() if which == "ref" else restated()
    ws, hs, dp, sp, cp, cm, keep = _fuse_args(maps, cams)
    fn = lib.ref_fuse if which == "ref" else lib.orc_fuse
    n = fn(len(maps), ws.ctypes.data, hs.ctypes.data, dp, sp, cp, cm, C.byref(cfg), 0, None, None,
           None, None)
    pos = np.zeros((max(n, 1), 3)); nrm = np.zeros((max(n, 1), 3))
    conf = np.zeros(max(n, 1), np.float32)
    extra = np.zeros((max(n, 1), 3), np.float32) if which == "ref" else np.zeros(max(n, 1), np.int64)
    fn(len(maps), ws.ctypes.data, hs.ctypes.data, dp, sp, cp, cm, C.byref(cfg), n,
       pos.ctypes.data, nrm.ctypes.data, conf.ctypes.data, extra.ctypes.data)
    out = dict(positions=pos[:n], normals=nrm[:n], confidences=conf[:n])
    if which == "ref":
        out["colors"] = extra[:n]
    else:
        out["keys"] = extra[:n]
    return out


def select_views(which, ref_cam, poses, cfg):
    lib = reference() if which == "ref" else restated()
    fn = lib.ref_select_views if which == "ref" else lib.orc_select_views
    out = np.zeros(max(len(poses), 1), np.int32)
    n = fn(C.byref(ref_cam), len(poses), _cams(poses), C.byref(cfg), out.ctypes.data, len(out))
    return out[:n].tolist()


def xorshift_stream(seed, n):
    out = np.zeros(n, np.uint32)
    restated().orc_xorshift_stream(seed, n, out.ctypes.data)
    return out


# --------------------------------------------------------------------------
# Plane mode (FP64)
# --------------------------------------------------------------------------

def _views(images, cams):
    imgs = [np.ascontiguousarray(i, np.float32) for i in images]
    return imgs, _ptr_array(imgs), _cams(cams)


def plane_eval(images, cams, ref, srcs, params, hyps, sample_shift=0.5):
    """FP64 cost of hypotheses rows (x, y, d, nx, ny, nz).  Returns (cost, nvalid, view_ncc)."""
    lib = restated()
    imgs, ptrs, cm = _views(images, cams)
    srcs = np.ascontiguousarray(srcs, np.int32)
    hyps = np.ascontiguousarray(hyps, np.float32).reshape(-1, 6)
    n = hyps.shape[0]
    cost = np.zeros(n); nv = np.zeros(n, np.int32); vn = np.zeros((n, len(srcs)))
    ok = lib.orc_plane_eval(len(imgs), ptrs, cm, ref, len(srcs), srcs.ctypes.data, C.byref(params),
                            sample_shift, n, hyps.ctypes.data, cost.ctypes.data, nv.ctypes.data,
                            vn.ctypes.data)
    assert ok
    return cost, nv, vn


def plane_patchmatch(images, cams, ref, srcs, params, init_planes=None):
    """FP64 plane-mode PatchMatch.  Returns (planes HxWx4 float32, cost HxW float64, support)."""
    lib = restated()
    imgs, ptrs, cm = _views(images, cams)
    srcs = np.ascontiguousarray(srcs, np.int32)
    h, w = imgs[ref].shape
    planes = np.zeros((h, w, 4), np.float32)
    cost = np.zeros((h, w)); sup = np.zeros((h, w), np.uint8)
    ip = None if init_planes is None else np.ascontiguousarray(init_planes, np.float32)
    ok = lib.orc_plane_patchmatch(len(imgs), ptrs, cm, ref, len(srcs), srcs.ctypes.data,
                                  C.byref(params), None if ip is None else ip.ctypes.data,
                                  planes.ctypes.data, cost.ctypes.data, sup.ctypes.data)
    assert ok
    return planes, cost, sup


def plane_refine(images, cams, ref, srcs, params, planes, cost, support):
    """FP64 plane-mode sub-pixel refinement (the search of the GPU's
    k_plane_refine) of a depth map: planes HxWx4 float32, cost HxW (>= 1e29:
    invalid), support HxW uint8.  Returns refined copies (planes, cost, support)."""
    lib = restated()
    imgs, ptrs, cm = _views(images, cams)
    srcs = np.ascontiguousarray(srcs, np.int32)
    planes = np.array(planes, np.float32, order="C")
    cost = np.array(cost, np.float64, order="C")
    support = np.array(support, np.uint8, order="C")
    ok = lib.orc_plane_refine(len(imgs), ptrs, cm, ref, len(srcs), srcs.ctypes.data,
                              C.byref(params), planes.ctypes.data, cost.ctypes.data,
                              support.ctypes.data)
    assert ok
    return planes, cost, support


def refine_1d_closed_form(kind, m, A, inv0, inv_per_px, iterations, inv_min, inv_max):
    """The refinement search on a closed-form cost (0: A (x - m)^2, 1: A |x - m|,
    2: 1 - exp(-((x - m) / A)^2)).  Returns (best x as float32, its cost,
    number of cost evaluations)."""
    bi = C.c_float()
    bc = C.c_double()
    n = restated().orc_refine_1d_closed_form(kind, m, A, inv0, inv_per_px, iterations, inv_min,
                                             inv_max, C.byref(bi), C.byref(bc))
    return np.float32(bi.value), bc.value, n


def seed_cells(w, h, seeds, grid=32):
    """Sparse-point seeding of PatchMatchStereo (core/src/mvs/patchmatch_stereo.cpp:
    499-533): seeds (u, v, depth) outside the image or with depth <= 0 are
    dropped; the rest fall into grid x grid pixel cells (column
    min(gx - 1, int(u) // grid)); a cell's depth is the mean of its seeds,
    summed in seed order.  Returns (gy, gx) float64, NaN for empty cells."""
    gx, gy = (w + grid - 1) // grid, (h + grid - 1) // grid
    sums = [[0.0] * gx for _ in range(gy)]
    counts = [[0] * gx for _ in range(gy)]
    for su, sv, sd in np.asarray(seeds, np.float64).reshape(-1, 3).tolist():
        if su < 0 or sv < 0 or su >= w or sv >= h or sd <= 0.0:
            continue
        cx, cy = min(gx - 1, int(su) // grid), min(gy - 1, int(sv) // grid)
        sums[cy][cx] += sd
        counts[cy][cx] += 1
    out = np.full((gy, gx), np.nan)
    for cy in range(gy):
        for cx in range(gx):
            if counts[cy][cx]:
                out[cy, cx] = sums[cy][cx] / counts[cy][cx]
    return out


def seeded_init_planes(w, h, seeds, grid=32):
    """Plane-mode seeded init: a pixel of a seeded cell starts on the
    fronto-parallel plane (normal (0, 0, -1)) at the cell's depth (float32),
    the reference's depth-only hypothesis; other rows are NaN (random init)."""
    cells = seed_cells(w, h, seeds, grid)
    ys, xs = np.mgrid[0:h, 0:w]
    d = cells[np.minimum(ys // grid, cells.shape[0] - 1), np.minimum(xs // grid, cells.shape[1] - 1)]
    planes = np.full((h, w, 4), np.nan, np.float32)
    ok = np.isfinite(d)
    planes[ok] = np.stack([d[ok].astype(np.float32), np.zeros(ok.sum(), np.float32),
                           np.zeros(ok.sum(), np.float32), np.full(ok.sum(), -1.0, np.float32)], 1)
    return planes


def geo_consistency(depths, cams, thr):
    lib = restated()
    ds = [np.ascontiguousarray(d, np.float32) for d in depths]
    outs = [np.zeros(d.shape, np.uint8) for d in ds]
    lib.orc_geo_consistency(len(ds), _ptr_array(ds), _cams(cams), thr, _ptr_array(outs))
    return outs


def so3_exp(w):
    """Rodrigues: rotation matrix of the rotation vector w (FP64)."""
    w = np.asarray(w, np.float64)
    th = float(np.linalg.norm(w))
    K = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    if th < 1e-12:
        return np.eye(3) + K + 0.5 * K @ K
    return np.eye(3) + (np.sin(th) / th) * K + ((1.0 - np.cos(th)) / (th * th)) * (K @ K)


def perturb_camera(cam, delta):
    """Copy of `cam` with its world->camera transform multiplied on the left by
    Exp(delta[:3]) and translated by delta[3:]: Y -> Exp(w) Y + tau."""
    import copy
    c = copy.copy(cam)
    R = np.array(cam.R[:], np.float64).reshape(3, 3)
    t = np.array(cam.t[:], np.float64)
    E = so3_exp(delta[:3])
    R2 = E @ R
    t2 = E @ t + np.asarray(delta[3:], np.float64)
    for i in range(9):
        c.R[i] = float(R2.reshape(-1)[i])
    for i in range(3):
        c.t[i] = float(t2[i])
    return c


def plane_cost_grad(images, cams, ref, srcs, params, xy, planes, weights=None, grad=True,
                    pose_grad=True):
    """FP64 point-form cost of planes (n x 4: D, nx, ny, nz, float64) at pixels xy
    (n x 2 int).  Returns (cost, nvalid, grad_plane n x 4 or None, grad_pose
    nsrc x 6 or None); gradients are of sum(weights * cost)."""
    lib = restated()
    imgs, ptrs, cm = _views(images, cams)
    srcs = np.ascontiguousarray(srcs, np.int32)
    xy = np.ascontiguousarray(xy, np.int32).reshape(-1, 2)
    planes = np.ascontiguousarray(planes, np.float64).reshape(-1, 4)
    n = xy.shape[0]
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    cost = np.zeros(n); nv = np.zeros(n, np.int32)
    gp = np.zeros((n, 4)) if grad else None
    gs = np.zeros((len(srcs), 6)) if (grad and pose_grad) else None
    ok = lib.orc_plane_cost_grad(len(imgs), ptrs, cm, ref, len(srcs), srcs.ctypes.data,
                                 C.byref(params), n, xy.ctypes.data, planes.ctypes.data,
                                 None if w is None else w.ctypes.data, cost.ctypes.data,
                                 nv.ctypes.data, None if gp is None else gp.ctypes.data,
                                 None if gs is None else gs.ctypes.data)
    assert ok
    return cost, nv, gp, gs


def bgra_to_gray(bgra):
    """BgraToGray (core/src/orchestrator/orchestrator.cpp:53-67) restated in
    numpy: gray = (77 R + 150 G + 29 B) >> 8 on B,G,R,A bytes."""
    a = np.asarray(bgra, np.int64)
    return ((77 * a[..., 2] + 150 * a[..., 1] + 29 * a[..., 0]) >> 8).astype(np.uint8)


# ---------------------------------------------------------------------------
# Downstream formats (host restatements; the reference has them as a spec or
# as plain C++ over std::vector, no kernels)
# ---------------------------------------------------------------------------


def to_oriented_cloud(positions, normals, colors, confidences):
    """UnifiedPointCloud::ToOrientedCloud (core/src/orchestrator/orchestrator.cpp:
    181-192): positions / normals / colors copied, weight = max(0, confidence),
    1.0 for points past the end of the confidence array."""
    n = len(positions)
    w = np.ones(n, np.float32)
    m = min(n, len(confidences))
    w[:m] = np.maximum(np.float32(0.0), np.asarray(confidences, np.float32)[:m])
    return (np.array(positions, np.float64), np.array(normals, np.float64),
            np.array(colors, np.float32), w)


SPILL_MAGIC = b"RMVD"


def spill_encode(ref_index, width, height, depth, support, confidence):
    """Bytes of mvs_depth/ref_%05d.f32 (docs/13_内存与降级.md §3.3.2: the depth
    map "with confidence and support counts"): a 32-byte header
    (magic, version 1, width, height, ref_index, 3 planes, 2 reserved) and the
    float32 planes depth, confidence, support."""
    hdr = SPILL_MAGIC + np.array([1], "<u4").tobytes() + \
        np.array([width, height, ref_index], "<i4").tobytes() + np.array([3, 0, 0], "<u4").tobytes()
    planes = np.concatenate([np.asarray(depth, "<f4").reshape(-1),
                             np.asarray(confidence, "<f4").reshape(-1),
                             np.asarray(support, np.uint8).reshape(-1).astype("<f4")])
    return hdr + planes.tobytes()


def spill_decode(blob):
    """Inverse of spill_encode: (ref_index, width, height, depth, support,
    confidence) or None when the bytes are not a well-formed map."""
    if len(blob) < 32 or blob[:4] != SPILL_MAGIC:
        return None
    ver = int(np.frombuffer(blob, "<u4", 1, 4)[0])
    w, h, ref = (int(x) for x in np.frombuffer(blob, "<i4", 3, 8))
    planes = int(np.frombuffer(blob, "<u4", 1, 20)[0])
    if ver != 1 or planes != 3 or w <= 0 or h <= 0 or len(blob) != 32 + 12 * w * h:
        return None
    f = np.frombuffer(blob, "<f4", 3 * w * h, 32)
    sup = f[2 * w * h:]
    if not np.all((sup >= 0) & (sup <= 255) & (sup == np.floor(sup))):
        return None
    return ref, w, h, f[:w * h].copy(), sup.astype(np.uint8), f[w * h:2 * w * h].copy()


def spill_manifest_item(ref_index, blob):
    """The manifest entry the spill records (§3.3.3: path, bytes, sha256, stage)."""
    import hashlib
    return {"path": "mvs_depth/ref_%05d.f32" % ref_index, "bytes": len(blob),
            "sha256": hashlib.sha256(blob).hexdigest(), "stage": "mvs"}
