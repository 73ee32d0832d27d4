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
