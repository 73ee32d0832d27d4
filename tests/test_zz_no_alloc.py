"""The allocation-free adjoint entry point (recon_mvs_hip_plane_cost_grad_noalloc,
include/recon_mvs_hip.h) against the allocating one it varies
(recon_mvs_hip_plane_cost_grad): bit-identical costs, plane gradients and
pose gradients, on two calls in a row, the second on a smaller map (the
shared scratch array keeps the first call's partials, which must not leak
into the second); and inputs beyond the scratch's 2 MiB are refused."""
import ctypes as C

import numpy as np
import pytest

from recon_scanner_amd import mvs, plane, scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


class DeviceBuffers:
    """Device buffers owned by the test, from the HIP runtime the library
    itself is linked to (libamdhip64.so.7; a second runtime in the process,
    such as torch's, would not see the device)."""

    def __init__(self):
        self.H = C.CDLL("libamdhip64.so.7")
        self.H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.H.hipFree.argtypes = [C.c_void_p]
        self.H.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def alloc(self, arr):
        """A device copy of the host array `arr`."""
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.H.hipMalloc(C.byref(p), arr.nbytes) == 0
        self.ptrs.append(p.value)
        assert self.H.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0
        return p.value

    def fetch(self, ptr, like):
        out = np.empty_like(like)
        assert self.H.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    def close(self):
        self.H.hipDeviceSynchronize()
        for q in self.ptrs:
            self.H.hipFree(C.c_void_p(q))
        self.ptrs = []


def _declare(L):
    v = C.c_void_p
    L.recon_mvs_hip_plane_cost_grad_noalloc.argtypes = [v, C.c_int32, C.c_int32, v,
                                                        C.POINTER(plane.PlaneParams), v, v, v, v, v]
    L.recon_mvs_hip_plane_cost_grad_noalloc.restype = C.c_int32


def _noalloc(pm, ref, srcs, p, pose, weights):
    L = mvs.lib()
    _declare(L)
    h, w = pm.shapes[ref]
    s = np.ascontiguousarray(srcs, np.int32)
    pd = np.ascontiguousarray(pose, np.float64)
    bufs = DeviceBuffers()
    try:
        dw = bufs.alloc(np.ascontiguousarray(weights, np.float32).reshape(-1))
        c = np.full((h, w), -7.0, np.float32)
        g = np.full((h, w, 4), -7.0, np.float32)
        sp = np.full((len(s), 6), -7.0, np.float64)
        dc, dg, dp = bufs.alloc(c), bufs.alloc(g), bufs.alloc(sp)
        st = L.recon_mvs_hip_plane_cost_grad_noalloc(pm.ctx, ref, len(s), s.ctypes.data, C.byref(p),
                                                     pd.ctypes.data, dw, dc, dg, dp)
        mvs.check(st, "plane_cost_grad_noalloc")
        mvs.check(L.recon_mvs_hip_synchronize(pm.ctx), "synchronize")
        return bufs.fetch(dc, c), bufs.fetch(dg, g), bufs.fetch(dp, sp)
    finally:
        bufs.close()


def test_noalloc_cost_grad_equals_allocating_entry_point():
    rng = np.random.default_rng(3)
    for w, h in ((480, 360), (240, 180)):     # the second call on a smaller map
        say("noalloc vs allocating cost_grad at", w, "x", h)
        sc = scenes.sphere_planes(width=w, height=h, focal=625.0 * w / 480.0, n_views=7, supersample=1)
        srcs = list(range(1, 7))
        p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=2,
                                 min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0)
        pm = plane.PlaneMVS(sc.cams, sc.images, storage="f32")
        pm.patchmatch(0, srcs, p)
        pose = np.concatenate([rng.normal(scale=2e-3, size=(6, 3)), rng.normal(scale=5e-4, size=(6, 3))], 1)
        weights = rng.uniform(0.5, 1.5, size=(h, w)).astype(np.float32)
        c0, g0, s0 = pm.cost_grad(0, srcs, p, pose_delta=pose, weights=weights)
        c1, g1, s1 = _noalloc(pm, 0, srcs, p, pose, weights)
        fin = c0 < 1e29
        say("valid pixels", int(fin.sum()), "pose grad norm", float(np.linalg.norm(s0)))
        assert fin.mean() > 0.5 and np.linalg.norm(s0) > 0
        assert c1.tobytes() == c0.tobytes()
        assert g1.tobytes() == g0.tobytes()
        assert s1.tobytes() == s0.tobytes()


def test_noalloc_refuses_partials_beyond_its_scratch():
    """1920x1440 (5400 tiles of 32 x 16) against 9 sources needs 2.33 MB of
    pose partials, above the 2 MiB scratch: refused, with the sizes named;
    the same map without the pose gradient needs no scratch and runs."""
    say("noalloc refusal at 1920x1440 with 9 sources")
    sc = scenes.sphere_planes(width=1920, height=1440, focal=2500.0, n_views=10, supersample=1)
    srcs = list(range(1, 10))
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=0,
                             min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0)
    pm = plane.PlaneMVS(sc.cams, sc.images, storage="f32")
    pm.patchmatch(0, srcs, p)
    L = mvs.lib()
    _declare(L)
    s = np.ascontiguousarray(srcs, np.int32)
    bufs = DeviceBuffers()
    try:
        c = np.zeros((1440, 1920), np.float32)
        dc, dp = bufs.alloc(c), bufs.alloc(np.zeros(6 * 9))
        st = L.recon_mvs_hip_plane_cost_grad_noalloc(pm.ctx, 0, 9, s.ctypes.data, C.byref(p), None,
                                                     None, dc, None, dp)
        msg = L.recon_mvs_hip_last_error(pm.ctx).decode()
        say("status", st, msg)
        assert st == 1 and "2 MiB" in msg and "5400 tiles" in msg
        st = L.recon_mvs_hip_plane_cost_grad_noalloc(pm.ctx, 0, 9, s.ctypes.data, C.byref(p), None,
                                                     None, dc, None, None)
        assert st == 0
        mvs.check(L.recon_mvs_hip_synchronize(pm.ctx), "synchronize")
        c0, _, _ = pm.cost_grad(0, srcs, p, grad=False)
        assert bufs.fetch(dc, c).tobytes() == c0.tobytes()
    finally:
        bufs.close()
