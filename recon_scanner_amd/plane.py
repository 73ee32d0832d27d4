"""Plane-mode PatchMatch on the GPU: views resident in HBM, slanted-plane
hypotheses with homography warps, top-k NCC, red-black sweeps, geometric
consistency (include/recon_mvs_hip.h, "Plane mode").

    pm = PlaneMVS(scene_cams, scene_images, storage="f32")
    pm.patchmatch(ref, srcs, params)        # asynchronous
    pm.refine(ref, srcs, params)            # optional sub-pixel pass
    pm.geo_consistency(views, 0.01)
    d = pm.depth_map(ref)                   # synchronous copy to the host

No CPU fallback: every call goes through librecon_mvs_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from . import mvs

FORMATS = {"f32": N.PIXEL_F32, "f16": N.PIXEL_F16, "bf16": N.PIXEL_BF16, "u8": N.PIXEL_U8}
KERNEL_CLASSES = ("sweep", "init", "refine", "consistency", "score", "eval", "batch", "adjoint")


class PlaneParams(C.Structure):
    _fields_ = [("patch_size", C.c_int32), ("window_stride", C.c_int32), ("top_k", C.c_int32),
                ("iterations", C.c_int32), ("seed", C.c_uint32), ("propagation_reach", C.c_int32),
                ("refine_iterations", C.c_int32), ("tile_schedule", C.c_int32),
                ("min_depth_m", C.c_double), ("max_depth_m", C.c_double),
                ("depth_perturb", C.c_double), ("normal_perturb", C.c_double),
                ("photo_threshold", C.c_double), ("value_scale", C.c_double),
                ("occupancy_cap", C.c_int32), ("capped_iterations", C.c_int32),
                ("view_phased_iterations", C.c_int32), ("binned_phases", C.c_int32)]

    def copy(self, **kw):
        p = PlaneParams()
        C.memmove(C.addressof(p), C.addressof(self), C.sizeof(p))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        return p


class BatchJob(C.Structure):
    _fields_ = [("ref_view", C.c_int32), ("num_src", C.c_int32), ("src_views", C.c_int32 * 16)]


class View(C.Structure):
    _fields_ = [("pixels", C.c_void_p), ("pixel_format", C.c_int32), ("row_stride", C.c_int32),
                ("camera", N.Camera)]


_declared = False


def _lib():
    global _declared
    L = mvs.lib()
    if not _declared:
        P = C.POINTER
        v = C.c_void_p
        L.recon_mvs_plane_params_default.argtypes = [P(PlaneParams)]
        L.recon_mvs_plane_params_default.restype = None
        L.recon_mvs_hip_set_views.argtypes = [v, C.c_int32, P(View), C.c_int32]
        L.recon_mvs_hip_plane_patchmatch.argtypes = [v, C.c_int32, C.c_int32, v, P(PlaneParams)]
        L.recon_mvs_hip_plane_patchmatch_seeded.argtypes = [v, C.c_int32, C.c_int32, v,
                                                            P(PlaneParams), v, C.c_int32]
        L.recon_mvs_hip_plane_refine.argtypes = [v, C.c_int32, C.c_int32, v, P(PlaneParams)]
        L.recon_mvs_hip_geo_consistency.argtypes = [v, C.c_int32, v, C.c_double, C.c_int32]
        L.recon_mvs_hip_get_depth_map.argtypes = [v, C.c_int32, C.c_int32, v, v, v, v, v, v]
        L.recon_mvs_hip_set_plane_state.argtypes = [v, C.c_int32, C.c_int32, v, P(PlaneParams), v, v]
        L.recon_mvs_hip_eval_plane_costs.argtypes = [v, C.c_int32, C.c_int32, v, P(PlaneParams),
                                                     C.c_int32, v, v, v, v]
        L.recon_mvs_hip_timing.argtypes = [v, C.c_int32, C.c_int32, v, v]
        L.recon_mvs_hip_autotune.argtypes = [v, C.c_int32, C.c_int32, v, P(PlaneParams), v, v]
        L.recon_mvs_hip_batch_create.argtypes = [v, C.c_int32, P(BatchJob), P(PlaneParams),
                                                 P(C.c_void_p)]
        L.recon_mvs_hip_batch_run.argtypes = [v, v]
        L.recon_mvs_hip_batch_schedule.argtypes = [v]
        L.recon_mvs_hip_batch_destroy.argtypes = [v, v]
        L.recon_mvs_hip_plane_cost_grad.argtypes = [v, C.c_int32, C.c_int32, v, P(PlaneParams),
                                                    v, v, v, v, v]
        L.recon_mvs_hip_plane_polish.argtypes = [v, C.c_int32, C.c_int32, v, P(PlaneParams),
                                                 C.c_int32]
        L.recon_mvs_hip_bgra_to_gray.argtypes = [v, v, C.c_int32, C.c_int32, C.c_int32, v]
        _declared = True
    return L


def default_params(**kw) -> PlaneParams:
    p = PlaneParams()
    _lib().recon_mvs_plane_params_default(C.byref(p))
    return p.copy(**kw)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


class PlaneMVS:
    """A set of views resident on the GPU and their depth-map slots."""

    def __init__(self, cams, images, storage="f32"):
        self.L = _lib()
        self.ctx = mvs.context()
        self.cams = list(cams)
        self.shapes = [(c.height, c.width) for c in self.cams]
        views = (View * len(self.cams))()
        keep = []
        for i, (c, img) in enumerate(zip(self.cams, images)):
            img = np.asarray(img)
            stride = 0
            if img.ndim == 3:
                # B,G,R,A frames (rows may be padded: any row stride >= 4 x width)
                assert img.dtype == np.uint8 and img.shape == (c.height, c.width, 4), img.shape
                a, fmt = img, N.PIXEL_BGRA8
                if not (a.strides[2] == 1 and a.strides[1] == 4 and a.strides[0] >= 4 * c.width):
                    a = np.ascontiguousarray(a)
                stride = a.strides[0]
            elif img.dtype == np.uint8:
                a, fmt = np.ascontiguousarray(img), N.PIXEL_U8
            else:
                a, fmt = np.ascontiguousarray(img, np.float32), N.PIXEL_F32
            if img.ndim == 2:
                assert a.shape == (c.height, c.width), (a.shape, c.height, c.width)
            keep.append(a)
            views[i].pixels = a.ctypes.data
            views[i].pixel_format = fmt
            views[i].row_stride = stride
            views[i].camera = c
        self._check(self.L.recon_mvs_hip_set_views(self.ctx, len(self.cams), views, FORMATS[storage]),
                    "set_views")
        self.storage = storage

    def _check(self, st, what):
        mvs.check(st, what)

    def patchmatch(self, ref, srcs, params, seed_points=None):
        """Plane-mode PatchMatch of view ref; seed_points: optional (n, 3)
        rows (u, v, depth_m), the sparse-point seeding of PatchMatchStereo
        (patchmatch_stereo.cpp:499-533)."""
        s = _i32(srcs)
        if seed_points is None:
            self._check(self.L.recon_mvs_hip_plane_patchmatch(self.ctx, ref, len(s), s.ctypes.data,
                                                              C.byref(params)), "plane_patchmatch")
            return
        sp = np.ascontiguousarray(seed_points, np.float64).reshape(-1, 3)
        self._check(self.L.recon_mvs_hip_plane_patchmatch_seeded(
            self.ctx, ref, len(s), s.ctypes.data, C.byref(params), sp.ctypes.data, sp.shape[0]),
            "plane_patchmatch_seeded")

    def refine(self, ref, srcs, params):
        s = _i32(srcs)
        self._check(self.L.recon_mvs_hip_plane_refine(self.ctx, ref, len(s), s.ctypes.data,
                                                      C.byref(params)), "plane_refine")

    def geo_consistency(self, views, threshold, min_views=2):
        v = _i32(views)
        self._check(self.L.recon_mvs_hip_geo_consistency(self.ctx, len(v), v.ctypes.data, threshold,
                                                         min_views), "geo_consistency")

    def set_state(self, ref, srcs, params, depth, normals):
        s = _i32(srcs)
        d = np.ascontiguousarray(depth, np.float32)
        n = np.ascontiguousarray(normals, np.float32)
        self._check(self.L.recon_mvs_hip_set_plane_state(self.ctx, ref, len(s), s.ctypes.data,
                                                         C.byref(params), d.ctypes.data,
                                                         n.ctypes.data), "set_plane_state")

    def synchronize(self):
        self._check(self.L.recon_mvs_hip_synchronize(self.ctx), "synchronize")

    def depth_map(self, view, drop_inconsistent=0, geo=False):
        h, w = self.shapes[view]
        depth = np.zeros((h, w), np.float32)
        normals = np.zeros((h, w, 3), np.float32)
        cost = np.zeros((h, w), np.float32)
        sup = np.zeros((h, w), np.uint8)
        conf = np.zeros((h, w), np.float32)
        gs = np.zeros((h, w), np.uint8) if geo else None
        self._check(self.L.recon_mvs_hip_get_depth_map(
            self.ctx, view, drop_inconsistent, depth.ctypes.data, normals.ctypes.data,
            cost.ctypes.data, sup.ctypes.data, conf.ctypes.data,
            None if gs is None else gs.ctypes.data), "get_depth_map")
        out = dict(depth=depth, normals=normals, cost=cost, support=sup, confidence=conf)
        if geo:
            out["geo_support"] = gs
        return out

    def eval_costs(self, ref, srcs, params, hyps):
        s = _i32(srcs)
        hyps = np.ascontiguousarray(hyps, np.float32).reshape(-1, 6)
        n = hyps.shape[0]
        cost = np.zeros(n, np.float32)
        nv = np.zeros(n, np.int32)
        vn = np.zeros((n, len(s)), np.float32)
        self._check(self.L.recon_mvs_hip_eval_plane_costs(self.ctx, ref, len(s), s.ctypes.data,
                                                          C.byref(params), n, hyps.ctypes.data,
                                                          cost.ctypes.data, nv.ctypes.data,
                                                          vn.ctypes.data), "eval_plane_costs")
        return cost, nv, vn

    def cost_grad(self, ref, srcs, params, pose_delta=None, weights=None, grad=True,
                  pose_grad=True):
        """Cost of every pixel of the slot of `ref` (under source poses perturbed
        by pose_delta, nsrc x 6: w, tau) and the gradients of sum(weights * cost):
        returns (cost HxW, grad_plane HxWx4 or None, grad_pose nsrc x 6 or None)."""
        s = _i32(srcs)
        h, w = self.shapes[ref]
        pd = None if pose_delta is None else np.ascontiguousarray(pose_delta, np.float64).reshape(len(s), 6)
        wt = None if weights is None else np.ascontiguousarray(weights, np.float32).reshape(h, w)
        cost = np.zeros((h, w), np.float32)
        gp = np.zeros((h, w, 4), np.float32) if grad else None
        gs = np.zeros((len(s), 6)) if (grad and pose_grad) else None
        self._check(self.L.recon_mvs_hip_plane_cost_grad(
            self.ctx, ref, len(s), s.ctypes.data, C.byref(params),
            None if pd is None else pd.ctypes.data, None if wt is None else wt.ctypes.data,
            cost.ctypes.data, None if gp is None else gp.ctypes.data,
            None if gs is None else gs.ctypes.data), "plane_cost_grad")
        return cost, gp, gs

    def polish(self, ref, srcs, params, steps=10):
        """Gradient polish of the slot of `ref` (costs never increase)."""
        s = _i32(srcs)
        self._check(self.L.recon_mvs_hip_plane_polish(self.ctx, ref, len(s), s.ctypes.data,
                                                      C.byref(params), int(steps)), "plane_polish")

    def autotune(self, ref, srcs, params):
        """Times every tile schedule on this view; returns (best, [ms per schedule])."""
        s = _i32(srcs)
        best = C.c_int32(0)
        ms = np.zeros(8)
        self._check(self.L.recon_mvs_hip_autotune(self.ctx, ref, len(s), s.ctypes.data,
                                                  C.byref(params), C.byref(best), ms.ctypes.data),
                    "autotune")
        return int(best.value), ms[:3].tolist()

    def batch(self, jobs, params):
        """A captured batch over jobs = [(ref, [srcs])]; see Batch."""
        return Batch(self, jobs, params)

    def timing(self, enable=True, reset=False):
        ms = np.zeros(8)
        cnt = np.zeros(8, np.int64)
        self._check(self.L.recon_mvs_hip_timing(self.ctx, int(enable), int(reset), ms.ctypes.data,
                                                cnt.ctypes.data), "timing")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_CLASSES)}


def bgra_to_gray(bgra):
    """Gray of a B,G,R,A frame (H x W x 4 uint8, rows may be padded) on the
    GPU: recon_mvs_hip_bgra_to_gray, the reference's BgraToGray."""
    a = np.asarray(bgra)
    assert a.dtype == np.uint8 and a.ndim == 3 and a.shape[2] == 4
    if not (a.strides[2] == 1 and a.strides[1] == 4):
        a = np.ascontiguousarray(a)
    h, w = a.shape[:2]
    out = np.zeros((h, w), np.uint8)
    mvs.check(_lib().recon_mvs_hip_bgra_to_gray(mvs.context(), a.ctypes.data, w, h, a.strides[0],
                                                out.ctypes.data), "bgra_to_gray")
    return out


def round_storage(img, storage):
    """The values the GPU stores for `img` in `storage` (round to nearest even)."""
    a = np.asarray(img, np.float32)
    if storage == "f32":
        return a.copy()
    if storage == "f16":
        return a.astype(np.float16).astype(np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


class Batch:
    """Many reference views of any size in one persistent launch per phase,
    captured once into a HIP graph (recon_mvs_hip_batch_*)."""

    def __init__(self, pm: PlaneMVS, jobs, params):
        self.pm = pm
        arr = (BatchJob * len(jobs))()
        for i, (r, srcs) in enumerate(jobs):
            arr[i].ref_view = int(r)
            arr[i].num_src = len(srcs)
            for k, v in enumerate(srcs):
                arr[i].src_views[k] = int(v)
        h = C.c_void_p()
        pm._check(pm.L.recon_mvs_hip_batch_create(pm.ctx, len(jobs), arr, C.byref(params),
                                                  C.byref(h)), "batch_create")
        self.h = h
        self.refs = [int(r) for r, _ in jobs]
        self.schedule = int(pm.L.recon_mvs_hip_batch_schedule(h))

    def run(self):
        self.pm._check(self.pm.L.recon_mvs_hip_batch_run(self.pm.ctx, self.h), "batch_run")

    def close(self):
        if self.h:
            self.pm.L.recon_mvs_hip_batch_destroy(self.pm.ctx, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
