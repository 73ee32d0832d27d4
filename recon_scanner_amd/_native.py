"""ctypes mirror of include/recon_mvs_hip.h and the loaders of the built libraries.

The product library is ``_lib/librecon_mvs_hip.so`` (HIP kernels for gfx950 +
the C ABI + the C++ ``recon::`` entry points).  ``_lib/libreconscene.so`` is the
synthetic-scene renderer used to make inputs.  Nothing here falls back to a CPU
path: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "_lib")
# RECON_MVS_HIP_LIB overrides the library path (deployments that install it elsewhere)
HIP_LIB = os.environ.get("RECON_MVS_HIP_LIB", os.path.join(LIB_DIR, "librecon_mvs_hip.so"))
SCENE_LIB = os.path.join(LIB_DIR, "libreconscene.so")

# status codes (recon_c_api.h:135-176)
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_OUT_OF_MEMORY = 2
ERR_NOT_IMPLEMENTED = 3
ERR_INTERNAL = 4
ERR_CANCELLED = 5
ERR_NO_DEPTH = 500
ERR_GPU_FAILURE = 502

PIXEL_U8, PIXEL_F32, PIXEL_F16, PIXEL_BF16, PIXEL_BGRA8 = 0, 1, 2, 3, 4


class Camera(C.Structure):
    _fields_ = [
        ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
        ("k1", C.c_double), ("k2", C.c_double), ("k3", C.c_double),
        ("p1", C.c_double), ("p2", C.c_double),
        ("width", C.c_int32), ("height", C.c_int32),
        ("R", C.c_double * 9), ("t", C.c_double * 3),
    ]

    @staticmethod
    def make(fx, fy, cx, cy, width, height, R=None, t=None, dist=(0, 0, 0, 0, 0)):
        c = Camera()
        c.fx, c.fy, c.cx, c.cy = fx, fy, cx, cy
        c.k1, c.k2, c.k3, c.p1, c.p2 = dist
        c.width, c.height = int(width), int(height)
        R = np.eye(3) if R is None else np.asarray(R, dtype=np.float64)
        t = np.zeros(3) if t is None else np.asarray(t, dtype=np.float64)
        for i in range(9):
            c.R[i] = float(R.reshape(-1)[i])
        for i in range(3):
            c.t[i] = float(t[i])
        return c

    def R_np(self):
        return np.array(self.R[:], dtype=np.float64).reshape(3, 3)

    def t_np(self):
        return np.array(self.t[:], dtype=np.float64)

    def K(self):
        return np.array([[self.fx, 0, self.cx], [0, self.fy, self.cy], [0, 0, 1.0]])

    def center(self):
        return -(self.R_np().T @ self.t_np())


class Config(C.Structure):
    _fields_ = [
        ("num_views", C.c_int32), ("ref_view_index", C.c_int32),
        ("ideal_angle_min_deg", C.c_double), ("ideal_angle_max_deg", C.c_double),
        ("max_view_scale_mismatch", C.c_double), ("view_angle_weight", C.c_double),
        ("view_texture_weight", C.c_double), ("view_spread_weight", C.c_double),
        ("view_scale_weight", C.c_double), ("reference_depth_m", C.c_double),
        ("patch_size", C.c_int32), ("iterations", C.c_int32),
        ("random_search_count", C.c_int32), ("propagation_directions", C.c_int32),
        ("random_search_radius_ratio", C.c_double), ("restart_ratio", C.c_double),
        ("propagation_enabled", C.c_int32), ("num_threads", C.c_int32),
        ("photo_threshold", C.c_double), ("depth_regularization", C.c_double),
        ("geo_consistency_threshold", C.c_double), ("weak_texture_variance", C.c_double),
        ("subpixel_iterations", C.c_int32), ("subpixel_fit_points", C.c_int32),
        ("subpixel_window_ratio", C.c_double),
        ("min_depth_m", C.c_double), ("max_depth_m", C.c_double),
        ("min_view_consistency", C.c_int32), ("reserved0", C.c_int32),
        ("fusion_voxel_size_m", C.c_double),
    ]


def config_default() -> Config:
    """MvsConfig defaults (core/include/recon/mvs.h:24-67)."""
    c = Config()
    c.num_views = 7
    c.ref_view_index = -1
    c.ideal_angle_min_deg = 8.0
    c.ideal_angle_max_deg = 35.0
    c.max_view_scale_mismatch = 0.15
    c.view_angle_weight = 1.0
    c.view_texture_weight = 0.8
    c.view_spread_weight = 0.6
    c.view_scale_weight = 0.4
    c.reference_depth_m = 0.40
    c.patch_size = 11
    c.iterations = 4
    c.random_search_count = 5
    c.random_search_radius_ratio = 0.30
    c.restart_ratio = 0.05
    c.propagation_directions = 4
    c.propagation_enabled = 1
    c.num_threads = 0
    c.photo_threshold = 1.0
    c.depth_regularization = 0.02
    c.geo_consistency_threshold = 0.02
    c.weak_texture_variance = 12.0
    c.subpixel_iterations = 4
    c.subpixel_fit_points = 5
    c.subpixel_window_ratio = 0.005
    c.min_depth_m = 0.10
    c.max_depth_m = 1.50
    c.min_view_consistency = 2
    c.fusion_voxel_size_m = 0.001
    return c


class NccResult(C.Structure):
    _fields_ = [("valid", C.c_int32), ("weak_texture", C.c_int32),
                ("valid_samples", C.c_int32), ("reserved0", C.c_int32), ("ncc", C.c_double)]


class CostResult(C.Structure):
    _fields_ = [("cost", C.c_double), ("ncc_sum", C.c_double), ("weight_sum", C.c_double),
                ("weak_texture_count", C.c_int32), ("valid_views", C.c_int32)]


class Problem(C.Structure):
    _fields_ = [
        ("ref_pixels", C.c_void_p), ("ref_w", C.c_int32), ("ref_h", C.c_int32),
        ("num_src", C.c_int32), ("pixel_format", C.c_int32),
        ("src_pixels", C.POINTER(C.c_void_p)), ("src_w", C.POINTER(C.c_int32)),
        ("src_h", C.POINTER(C.c_int32)),
        ("ref_camera", Camera), ("src_cameras", C.POINTER(Camera)),
    ]


class DepthMapDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32),
                ("depth", C.POINTER(C.c_float)), ("support_count", C.POINTER(C.c_uint8)),
                ("confidence", C.POINTER(C.c_float))]


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_spheres", C.c_int32), ("num_planes", C.c_int32), ("num_boxes", C.c_int32),
        ("supersample", C.c_int32),
        ("spheres", (C.c_double * 4) * 8), ("planes", (C.c_double * 6) * 8),
        ("boxes", (C.c_double * 6) * 8),
        ("texture_freq", C.c_double), ("octaves", C.c_int32), ("seed", C.c_uint32),
        ("contrast", C.c_double), ("noise_sigma", C.c_double),
        ("noise_seed", C.c_uint32), ("reserved", C.c_int32),
    ]


def ptr(a: np.ndarray, ctype=None):
    """Pointer to a contiguous numpy array's data."""
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    if ctype is None:
        return C.c_void_p(a.ctypes.data)
    return a.ctypes.data_as(C.POINTER(ctype))


class ProblemHolder:
    """Builds a Problem struct and keeps the arrays it points at alive."""

    def __init__(self, ref_img: np.ndarray, src_imgs, ref_cam: Camera, src_cams, pixel_format=PIXEL_U8):
        self.ref = np.ascontiguousarray(ref_img)
        self.srcs = [np.ascontiguousarray(s) for s in src_imgs]
        n = len(self.srcs)
        self.src_ptr_arr = (C.c_void_p * max(n, 1))(*[s.ctypes.data for s in self.srcs])
        self.sw = np.array([s.shape[1] for s in self.srcs] or [0], dtype=np.int32)
        self.sh = np.array([s.shape[0] for s in self.srcs] or [0], dtype=np.int32)
        self.cams = (Camera * max(n, 1))(*src_cams)
        p = Problem()
        p.ref_pixels = self.ref.ctypes.data
        p.ref_h, p.ref_w = int(self.ref.shape[0]), int(self.ref.shape[1])
        p.num_src = n
        p.pixel_format = pixel_format
        p.src_pixels = C.cast(self.src_ptr_arr, C.POINTER(C.c_void_p))
        p.src_w = ptr(self.sw, C.c_int32)
        p.src_h = ptr(self.sh, C.c_int32)
        p.ref_camera = ref_cam
        p.src_cameras = C.cast(self.cams, C.POINTER(Camera))
        self.problem = p


_scene_lib = None


def scene_lib():
    global _scene_lib
    if _scene_lib is None:
        if not os.path.exists(SCENE_LIB):
            raise RuntimeError(f"{SCENE_LIB} is not built; run __graft_entry__.build()")
        lib = C.CDLL(SCENE_LIB)
        lib.recon_scene_render.argtypes = [C.POINTER(SceneDesc), C.POINTER(Camera),
                                           C.POINTER(C.c_float), C.POINTER(C.c_float)]
        lib.recon_scene_render.restype = C.c_int32
        lib.recon_scene_reference_plane.argtypes = [C.c_int32, C.c_double, C.POINTER(Camera),
                                                    C.POINTER(C.c_uint8), C.POINTER(C.c_float)]
        lib.recon_scene_reference_plane.restype = C.c_int32
        _scene_lib = lib
    return _scene_lib
