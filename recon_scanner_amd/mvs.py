"""Python mirror of the reference's MVS interface (core/include/recon/mvs.h:133-214),
implemented on the C ABI of the MI355X library (include/recon_mvs_hip.h).

Names, argument meaning and error behaviour follow the reference:
  * ``PatchMatchStereo`` / ``RefineSubpixel`` return the filled / refined pixel
    count and return 0 on the same invalid inputs (patchmatch_stereo.cpp:434-485,
    780-803), including the global-vs-relative intrinsics index contract
    (mvs.h:151-161);
  * ``FuseDepthMaps`` returns an empty cloud on empty input or voxel <= 0.
A GPU failure raises ``RuntimeError`` (the C++ layer maps it to
RECON_ERR_MVS_GPU_FAILURE = 502, recon_c_api.h:176).  There is no CPU fallback:
without the built HIP library every call raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import threading

import numpy as np

from . import _native as N

# ---------------------------------------------------------------------------
# Types (mvs.h:24-125, camera.h:24-78)
# ---------------------------------------------------------------------------


def MvsConfig(**kw) -> N.Config:
    """MvsConfig with the defaults of mvs.h:24-67; keyword overrides."""
    c = N.config_default()
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(k)
        setattr(c, k, v)
    return c


class CameraIntrinsics:
    def __init__(self, fx=1.0, fy=1.0, cx=0.0, cy=0.0, k1=0.0, k2=0.0, k3=0.0, p1=0.0, p2=0.0,
                 width=0, height=0):
        self.fx, self.fy, self.cx, self.cy = fx, fy, cx, cy
        self.k1, self.k2, self.k3, self.p1, self.p2 = k1, k2, k3, p1, p2
        self.width, self.height = width, height

    @staticmethod
    def from_camera(c: N.Camera):
        return CameraIntrinsics(c.fx, c.fy, c.cx, c.cy, c.k1, c.k2, c.k3, c.p1, c.p2, c.width,
                                c.height)


class CameraPose:
    """World -> camera (R, t)."""

    def __init__(self, R=None, t=None):
        self.R = np.eye(3) if R is None else np.asarray(R, np.float64).reshape(3, 3)
        self.t = np.zeros(3) if t is None else np.asarray(t, np.float64).reshape(3)

    @staticmethod
    def from_camera(c: N.Camera):
        return CameraPose(c.R_np(), c.t_np())

    def CenterWorld(self):
        return -(self.R.T @ self.t)


def _camera(pose: CameraPose, intr: CameraIntrinsics) -> N.Camera:
    return N.Camera.make(intr.fx, intr.fy, intr.cx, intr.cy, intr.width, intr.height, pose.R,
                         pose.t, (intr.k1, intr.k2, intr.k3, intr.p1, intr.p2))


class NccResult:
    def __init__(self, valid=False, ncc=0.0, weak_texture=False, valid_samples=0):
        self.valid, self.ncc, self.weak_texture, self.valid_samples = valid, ncc, weak_texture, valid_samples


class DepthMap:
    def __init__(self):
        self.width = 0
        self.height = 0
        self.depth = np.zeros(0, np.float32)
        self.support_count = np.zeros(0, np.uint8)
        self.confidence = np.zeros(0, np.float32)

    def Allocate(self, w, h):
        self.width, self.height = w, h
        self.depth = np.full(w * h, np.nan, np.float32)
        self.support_count = np.zeros(w * h, np.uint8)
        self.confidence = np.zeros(w * h, np.float32)

    def IsValid(self, x, y):
        if x < 0 or y < 0 or x >= self.width or y >= self.height:
            return False
        return bool(np.isfinite(self.depth[y * self.width + x]))


class PointCloud:
    def __init__(self):
        self.positions = np.zeros((0, 3))
        self.normals = np.zeros((0, 3))
        self.colors = np.zeros(0, np.float32)
        self.confidences = np.zeros(0, np.float32)
        self.voxel_keys = np.zeros(0, np.int64)

    def size(self):
        return int(self.positions.shape[0])


# ---------------------------------------------------------------------------
# Library handle
# ---------------------------------------------------------------------------

_lib = None
_ctx = None
_lock = threading.Lock()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(N.HIP_LIB):
            raise RuntimeError(f"{N.HIP_LIB} is not built; run __graft_entry__.build()")
        L = C.CDLL(N.HIP_LIB)
        P = C.POINTER
        ctxp = C.c_void_p
        L.recon_mvs_hip_create.argtypes = [C.c_int32, P(C.c_void_p)]
        L.recon_mvs_hip_destroy.argtypes = [ctxp]
        L.recon_mvs_hip_last_error.argtypes = [ctxp]
        L.recon_mvs_hip_last_error.restype = C.c_char_p
        L.recon_mvs_hip_synchronize.argtypes = [ctxp]
        L.recon_mvs_hip_last_status.argtypes = [ctxp]
        L.recon_mvs_hip_cancel.argtypes = [ctxp]
        L.recon_mvs_hip_is_cancelled.argtypes = [ctxp]
        L.recon_mvs_hip_clear_cancel.argtypes = [ctxp]
        L.recon_mvs_hip_set_cancel_callback.argtypes = [ctxp, CANCEL_FN, C.c_void_p]
        L.recon_mvs_hip_timing.argtypes = [ctxp, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
        L.recon_mvs_config_default.argtypes = [P(N.Config)]
        L.recon_mvs_config_default.restype = None
        L.recon_mvs_hip_compute_ncc.argtypes = [ctxp, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                                C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                                C.c_int32, P(N.NccResult)]
        L.recon_mvs_hip_evaluate_candidates.argtypes = [ctxp, P(N.Problem), P(N.Config), C.c_int32,
                                                        C.c_void_p, P(N.CostResult)]
        L.recon_mvs_hip_patchmatch_stereo.argtypes = [ctxp, P(N.Problem), P(N.Config), C.c_void_p,
                                                      C.c_int32, C.c_void_p, C.c_void_p,
                                                      C.c_void_p, P(C.c_uint32)]
        L.recon_mvs_hip_refine_subpixel.argtypes = [ctxp, P(N.Problem), P(N.Config), C.c_void_p,
                                                    C.c_void_p, P(C.c_uint32)]
        L.recon_mvs_hip_fuse_depth_maps.argtypes = [ctxp, C.c_int32, P(N.DepthMapDesc),
                                                    P(N.Camera), P(N.Config), P(C.c_int64)]
        L.recon_mvs_hip_fetch_point_cloud.argtypes = [ctxp, C.c_int64, C.c_void_p, C.c_void_p,
                                                      C.c_void_p, C.c_void_p, C.c_void_p]
        L.recon_mvs_hip_run_mvs_stage.argtypes = [ctxp, C.c_int32, C.c_void_p, C.c_void_p,
                                                  P(N.Config), C.c_int32, C.c_void_p, C.c_void_p,
                                                  C.c_void_p, C.c_void_p, P(C.c_int32),
                                                  P(C.c_int64)]
        L.recon_mvs_hip_fetch_oriented_cloud.argtypes = [ctxp, C.c_int64, C.c_void_p, C.c_void_p,
                                                         C.c_void_p, C.c_void_p, P(C.c_int64)]
        L.recon_mvs_spill_depth_map.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32,
                                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p]
        L.recon_mvs_load_depth_map.argtypes = [C.c_char_p, C.c_int32, C.c_int64, P(C.c_int32),
                                               P(C.c_int32), C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


CANCEL_FN = C.CFUNCTYPE(C.c_int32, C.c_void_p)


def Cancel():
    """recon_mvs_hip_cancel on the process-wide context: thread-safe; running
    and later calls return status 5 (cancelled) at their next check point."""
    check(lib().recon_mvs_hip_cancel(context()), "cancel")


def IsCancelled() -> bool:
    return bool(lib().recon_mvs_hip_is_cancelled(context()))


def ClearCancel():
    check(lib().recon_mvs_hip_clear_cancel(context()), "clear_cancel")


_cancel_cb = None


def SetCancelCallback(fn):
    """fn() -> bool polled at the library's check points (None removes it)."""
    global _cancel_cb
    _cancel_cb = None if fn is None else CANCEL_FN(lambda _u: 1 if fn() else 0)
    check(lib().recon_mvs_hip_set_cancel_callback(context(), _cancel_cb or CANCEL_FN(), None),
          "set_cancel_callback")


def LastStatus() -> int:
    """Status of the last failed call on the process-wide context (0: none)."""
    return int(lib().recon_mvs_hip_last_status(context()))


def context():
    """The process-wide context on HIP device 0 (created on first use)."""
    global _ctx
    with _lock:
        if _ctx is None:
            h = C.c_void_p()
            st = lib().recon_mvs_hip_create(0, C.byref(h))
            if st != N.OK:
                raise RuntimeError(f"recon_mvs_hip_create failed with status {st}")
            _ctx = h
    return _ctx


def check(st, what=""):
    if st != N.OK:
        msg = lib().recon_mvs_hip_last_error(context()).decode(errors="replace")
        raise RuntimeError(f"{what} failed with status {st}: {msg}")


# ---------------------------------------------------------------------------
# Entry points
# ---------------------------------------------------------------------------


def ComputeNcc(ref_img, ref_w, ref_h, src_img, src_w, src_h, ref_u, ref_v, src_u, src_v,
               patch_size) -> NccResult:
    """ComputeNcc (patchmatch_stereo.cpp:214-274), one query."""
    r = ComputeNccBatch(ref_img, ref_w, ref_h, src_img, src_w, src_h,
                        np.array([[ref_u, ref_v, src_u, src_v]], np.float64), patch_size)
    return r[0]


def ComputeNccBatch(ref_img, ref_w, ref_h, src_img, src_w, src_h, coords, patch_size):
    a = np.ascontiguousarray(ref_img, np.uint8).reshape(-1)
    b = np.ascontiguousarray(src_img, np.uint8).reshape(-1)
    coords = np.ascontiguousarray(coords, np.float64).reshape(-1, 4)
    n = coords.shape[0]
    out = (N.NccResult * max(n, 1))()
    check(lib().recon_mvs_hip_compute_ncc(context(), a.ctypes.data, ref_w, ref_h, b.ctypes.data,
                                          src_w, src_h, n, coords.ctypes.data, patch_size, out),
          "ComputeNcc")
    return [NccResult(bool(o.valid), o.ncc, bool(o.weak_texture), o.valid_samples) for o in out[:n]]


def _resolve(src_gray, src_w, src_h, src_poses, src_intrinsics, source_indices):
    """Index contract of mvs.h:151-161 / patchmatch_stereo.cpp:450-468.
    Returns the per-source cameras, or None where the reference returns 0."""
    n = len(source_indices)
    if len(src_gray) != n or len(src_w) != n or len(src_h) != n:
        return None
    glob = len(src_intrinsics) == len(src_poses)
    rel = len(src_intrinsics) == n
    if not glob and not rel:
        return None
    if glob:
        for si in source_indices:
            if si < 0 or si >= len(src_poses):
                return None
    cams = []
    for k, si in enumerate(source_indices):
        intr = src_intrinsics[si] if glob else src_intrinsics[k]
        cams.append(_camera(src_poses[si], intr))
    return cams


def _problem(ref_gray, ref_w, ref_h, src_gray, src_w, src_h, ref_cam, src_cams):
    ref = np.ascontiguousarray(ref_gray, np.uint8).reshape(ref_h, ref_w)
    srcs = [np.ascontiguousarray(s, np.uint8).reshape(h, w) for s, w, h in zip(src_gray, src_w, src_h)]
    return N.ProblemHolder(ref, srcs, ref_cam, src_cams, N.PIXEL_U8)


def PatchMatchStereo(ref_gray, ref_w, ref_h, src_gray, src_w, src_h, ref_pose, src_poses, ref_intr,
                     src_intrinsics, source_indices, config, seed_points, out_depth: DepthMap) -> int:
    """PatchMatchStereo (mvs.h:193-206) on the GPU, reference mode."""
    if ref_gray is None or out_depth is None or len(source_indices) == 0:
        return 0
    cams = _resolve(src_gray, src_w, src_h, src_poses, src_intrinsics, source_indices)
    if cams is None:
        return 0
    holder = _problem(ref_gray, ref_w, ref_h, src_gray, src_w, src_h, _camera(ref_pose, ref_intr), cams)
    out_depth.Allocate(ref_w, ref_h)
    seeds = None
    if seed_points is not None and len(seed_points) > 0:
        seeds = np.ascontiguousarray(seed_points, np.float64).reshape(-1, 3)
    filled = C.c_uint32(0)
    check(lib().recon_mvs_hip_patchmatch_stereo(
        context(), C.byref(holder.problem), C.byref(config),
        None if seeds is None else seeds.ctypes.data, 0 if seeds is None else seeds.shape[0],
        out_depth.depth.ctypes.data, out_depth.support_count.ctypes.data,
        out_depth.confidence.ctypes.data, C.byref(filled)), "PatchMatchStereo")
    return int(filled.value)


def RefineSubpixel(ref_gray, ref_w, ref_h, src_gray, src_w, src_h, ref_pose, src_poses, ref_intr,
                   src_intrinsics, source_indices, config, depth: DepthMap) -> int:
    """RefineSubpixel (mvs.h:208-218) on the GPU, reference mode."""
    if ref_gray is None or depth is None or len(source_indices) == 0:
        return 0
    cams = _resolve(src_gray, src_w, src_h, src_poses, src_intrinsics, source_indices)
    if cams is None:
        return 0
    holder = _problem(ref_gray, ref_w, ref_h, src_gray, src_w, src_h, _camera(ref_pose, ref_intr), cams)
    refined = C.c_uint32(0)
    check(lib().recon_mvs_hip_refine_subpixel(context(), C.byref(holder.problem), C.byref(config),
                                              depth.depth.ctypes.data, depth.confidence.ctypes.data,
                                              C.byref(refined)), "RefineSubpixel")
    return int(refined.value)


def FuseDepthMaps(depth_maps, poses, intrinsics, config) -> PointCloud:
    """FuseDepthMaps (mvs.h:213-218) on the GPU.  Points in ascending voxel-key order."""
    cloud = PointCloud()
    if len(depth_maps) == 0 or config.fusion_voxel_size_m <= 0.0:
        return cloud
    descs = (N.DepthMapDesc * len(depth_maps))()
    cams = (N.Camera * len(depth_maps))()
    keep = []
    for i, dm in enumerate(depth_maps):
        cams[i] = _camera(poses[i], intrinsics[i])
        if dm is None:
            continue
        d = np.ascontiguousarray(dm.depth, np.float32)
        s = np.ascontiguousarray(dm.support_count, np.uint8)
        c = np.ascontiguousarray(dm.confidence, np.float32)
        keep += [d, s, c]
        descs[i].width, descs[i].height = dm.width, dm.height
        descs[i].depth = N.ptr(d, C.c_float)
        descs[i].support_count = N.ptr(s, C.c_uint8)
        descs[i].confidence = N.ptr(c, C.c_float)
    n = C.c_int64(0)
    check(lib().recon_mvs_hip_fuse_depth_maps(context(), len(depth_maps), descs, cams,
                                              C.byref(config), C.byref(n)), "FuseDepthMaps")
    m = int(n.value)
    cloud.positions = np.zeros((m, 3))
    cloud.normals = np.zeros((m, 3))
    cloud.colors = np.zeros(3 * m, np.float32)
    cloud.confidences = np.zeros(m, np.float32)
    cloud.voxel_keys = np.zeros(m, np.int64)
    if m:
        check(lib().recon_mvs_hip_fetch_point_cloud(context(), m, cloud.positions.ctypes.data,
                                                    cloud.normals.ctypes.data,
                                                    cloud.colors.ctypes.data,
                                                    cloud.confidences.ctypes.data,
                                                    cloud.voxel_keys.ctypes.data), "fetch")
    return cloud


STAGE_REFERENCE, STAGE_PLANE = 0, 1


def RunMvsStage(gray, cams, config, mode=STAGE_REFERENCE, plane_params=None):
    """The MVS stage over keyframes (Orchestrator::StageMvs, orchestrator.cpp:
    942-1071; recon_mvs_hip_run_mvs_stage).  gray: list of uint8 images or
    None; cams: list of N.Camera (pose + intrinsics).  Returns (produced,
    [DepthMap or None per keyframe], PointCloud)."""
    n = len(gray)
    imgs = [None if g is None else np.ascontiguousarray(g, np.uint8) for g in gray]
    gp = (C.c_void_p * max(n, 1))(*[None if g is None else g.ctypes.data for g in imgs])
    cams_arr = (N.Camera * max(n, 1))(*cams)
    maps = []
    dp = (C.c_void_p * max(n, 1))()
    sp = (C.c_void_p * max(n, 1))()
    cp = (C.c_void_p * max(n, 1))()
    for k in range(n):
        if imgs[k] is None:
            maps.append(None)
            continue
        dm = DepthMap()
        dm.Allocate(cams[k].width, cams[k].height)
        maps.append(dm)
        dp[k], sp[k], cp[k] = dm.depth.ctypes.data, dm.support_count.ctypes.data, dm.confidence.ctypes.data
    produced = C.c_int32(0)
    npts = C.c_int64(0)
    check(lib().recon_mvs_hip_run_mvs_stage(context(), n, gp, cams_arr, C.byref(config), mode,
                                            None if plane_params is None else C.byref(plane_params),
                                            dp, sp, cp, C.byref(produced), C.byref(npts)),
          "RunMvsStage")
    m = int(npts.value)
    cloud = PointCloud()
    cloud.positions = np.zeros((m, 3))
    cloud.normals = np.zeros((m, 3))
    cloud.colors = np.zeros(3 * m, np.float32)
    cloud.confidences = np.zeros(m, np.float32)
    cloud.voxel_keys = np.zeros(m, np.int64)
    if m:
        check(lib().recon_mvs_hip_fetch_point_cloud(context(), m, cloud.positions.ctypes.data,
                                                    cloud.normals.ctypes.data,
                                                    cloud.colors.ctypes.data,
                                                    cloud.confidences.ctypes.data,
                                                    cloud.voxel_keys.ctypes.data), "fetch")
    maps = [None if (d is None or not np.isfinite(d.depth).any()) else d for d in maps]
    return int(produced.value), maps, cloud


class OrientedPointCloud:
    """Poisson input (OrientedPointCloud, poisson_mesh.h:25-38)."""

    def __init__(self, n=0):
        self.positions = np.zeros((n, 3))
        self.normals = np.zeros((n, 3))
        self.colors = np.zeros(3 * n, np.float32)
        self.weights = np.zeros(n, np.float32)

    def size(self):
        return int(self.positions.shape[0])


def FetchOrientedCloud() -> OrientedPointCloud:
    """The cloud of the last FuseDepthMaps / RunMvsStage on the context in
    Poisson input form (UnifiedPointCloud::ToOrientedCloud, orchestrator.cpp:
    181-192): weights = max(0, confidence)."""
    n = C.c_int64(0)
    check(lib().recon_mvs_hip_fetch_oriented_cloud(context(), 0, None, None, None, None, C.byref(n)),
          "FetchOrientedCloud")
    out = OrientedPointCloud(int(n.value))
    if out.size():
        check(lib().recon_mvs_hip_fetch_oriented_cloud(context(), out.size(), out.positions.ctypes.data,
                                                       out.normals.ctypes.data, out.colors.ctypes.data,
                                                       out.weights.ctypes.data, None),
              "FetchOrientedCloud")
    return out


def SpillDepthMap(cache_dir, ref_index, dm: DepthMap) -> str:
    """Writes dm as <cache_dir>/mvs_depth/ref_%05d.f32 and records it in
    <cache_dir>/cache_manifest.json (docs/13_内存与降级.md §3.3.2-3.3.4);
    returns the file's sha256."""
    d = np.ascontiguousarray(dm.depth, np.float32).reshape(-1)
    s_ = np.ascontiguousarray(dm.support_count, np.uint8).reshape(-1)
    c = np.ascontiguousarray(dm.confidence, np.float32).reshape(-1)
    n = dm.width * dm.height
    if n <= 0 or d.size != n or s_.size != n or c.size != n:
        raise ValueError("SpillDepthMap: map buffers do not match width x height")
    hexd = C.create_string_buffer(65)
    st = lib().recon_mvs_spill_depth_map(os.fsencode(cache_dir), int(ref_index), dm.width, dm.height,
                                         d.ctypes.data, s_.ctypes.data, c.ctypes.data, hexd)
    if st != N.OK:
        raise RuntimeError(f"SpillDepthMap failed with status {st}")
    return hexd.value.decode()


def LoadDepthMap(cache_dir, ref_index):
    """Reads a spilled map back; None when it is missing or fails verification
    (the failed item is deleted from the cache: recompute it)."""
    w, h = C.c_int32(0), C.c_int32(0)
    path = os.fsencode(cache_dir)
    st = lib().recon_mvs_load_depth_map(path, int(ref_index), 0, C.byref(w), C.byref(h), None, None,
                                        None)
    if st == N.ERR_NO_DEPTH:
        return None
    if st != N.OK:
        raise RuntimeError(f"LoadDepthMap failed with status {st}")
    dm = DepthMap()
    dm.Allocate(w.value, h.value)
    st = lib().recon_mvs_load_depth_map(path, int(ref_index), w.value * h.value, C.byref(w),
                                        C.byref(h), dm.depth.ctypes.data,
                                        dm.support_count.ctypes.data, dm.confidence.ctypes.data)
    if st == N.ERR_NO_DEPTH:
        return None
    if st != N.OK:
        raise RuntimeError(f"LoadDepthMap failed with status {st}")
    return dm


def EvaluateCandidates(ref_gray, src_gray, ref_cam: N.Camera, src_cams, config, u_v_inv):
    """Batched EvaluateCandidate (patchmatch_stereo.cpp:344-418); returns CostResult list."""
    ref = np.ascontiguousarray(ref_gray, np.uint8)
    holder = N.ProblemHolder(ref, [np.ascontiguousarray(s, np.uint8) for s in src_gray], ref_cam,
                             src_cams, N.PIXEL_U8)
    q = np.ascontiguousarray(u_v_inv, np.float64).reshape(-1, 3)
    out = (N.CostResult * max(q.shape[0], 1))()
    check(lib().recon_mvs_hip_evaluate_candidates(context(), C.byref(holder.problem),
                                                  C.byref(config), q.shape[0], q.ctypes.data, out),
          "EvaluateCandidates")
    return out[:q.shape[0]]
