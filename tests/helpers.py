"""Shared test helpers: fixture loading and bit-level comparisons."""
import ctypes as C
import os

import numpy as np

from recon_scanner_amd import _native as N

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def cams_from_array(arr):
    out = []
    for r in arr:
        c = N.Camera.make(r[0], r[1], r[2], r[3], int(r[9]), int(r[10]), r[11:20].reshape(3, 3),
                          r[20:23], tuple(r[4:9]))
        out.append(c)
    return out


def cfg_from_bytes(b):
    cfg = N.Config()
    C.memmove(C.addressof(cfg), bytes(np.asarray(b, np.uint8)), C.sizeof(cfg))
    return cfg


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    return a.tobytes() == b.tobytes()


def sort_cloud(pos, nrm, conf):
    o = np.lexsort(np.asarray(pos).T[::-1])
    return np.asarray(pos)[o], np.asarray(nrm)[o], np.asarray(conf)[o]


def sphere_normals(sc, view=0):
    """Analytic normals of the sphere_arc scene in the camera frame of `view`."""
    cam = sc.cams[view]
    h, w = sc.depths[view].shape
    yy, xx = np.mgrid[0:h, 0:w]
    ray = np.stack([(xx + 0.5 - cam.cx) / cam.fx, (yy + 0.5 - cam.cy) / cam.fy, np.ones((h, w))], -1)
    Pc = ray * sc.depths[view][..., None]
    R, t = cam.R_np(), cam.t_np()
    Pw = (Pc - t) @ R                       # R^T (p - t)
    on_sphere = np.abs(np.linalg.norm(Pw - [0, 0, 0.35], axis=-1) - 0.05) < 1e-4
    nw = np.where(on_sphere[..., None], (Pw - [0, 0, 0.35]) / 0.05, [0.0, 0.0, -1.0])
    nc = nw @ R.T
    return nc.astype(np.float32)


class OracleBackend:
    """Plane-cost backend of the torch harness (recon_scanner_amd.adjoint) on
    the FP64 oracle: planes (n, 4) at pixels xy (n, 2) of view `ref`."""

    def __init__(self, sc, ref, srcs, params, xy):
        self.sc, self.ref, self.srcs, self.params = sc, ref, list(srcs), params
        self.xy = np.asarray(xy).reshape(-1, 2)

    def cams(self, pose_delta):
        import oracle
        cams = list(self.sc.cams)
        for k, v in enumerate(self.srcs):
            cams[v] = oracle.perturb_camera(self.sc.cams[v], pose_delta[k])
        return cams

    def evaluate(self, planes, pose_delta, weights=None, grad=False):
        import oracle
        planes = np.asarray(planes, np.float64)
        w = None if weights is None else np.asarray(weights, np.float64).reshape(-1)
        c, _, gp, gs = oracle.plane_cost_grad(self.sc.images, self.cams(pose_delta), self.ref,
                                              self.srcs, self.params, self.xy,
                                              planes.reshape(-1, 4), weights=w, grad=grad,
                                              pose_grad=grad)
        return c.reshape(planes.shape[:-1]), None if gp is None else gp.reshape(planes.shape), gs
