"""Writes the golden fixtures of the reference path (tests/golden/*.npz).

Runs the reference's own CPU code (compiled in place from core/src/mvs/
patchmatch_stereo.cpp, core/src/util/camera.cpp, core/src/util/linalg.cpp by
oracle/Makefile) on small seeded inputs and stores inputs and outputs as
plain arrays (np.savez, no pickles).  Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from recon_scanner_amd import _native as N  # noqa: E402
from recon_scanner_amd import scenes  # noqa: E402


def cams_array(cams):
    """Camera structs -> float64 rows: fx fy cx cy k1 k2 k3 p1 p2 w h R(9) t(3)."""
    rows = []
    for c in cams:
        rows.append([c.fx, c.fy, c.cx, c.cy, c.k1, c.k2, c.k3, c.p1, c.p2, c.width, c.height,
                     *c.R[:], *c.t[:]])
    return np.array(rows, np.float64)


def cfg_array(cfg):
    return np.frombuffer(bytes(cfg), np.uint8).copy()


def main():
    oracle.build()
    rng = np.random.default_rng(1234)

    # 1. ComputeNcc on a 64x64 texture: in-bounds, partial, out-of-bounds, flat
    sc = scenes.sphere_arc(width=64, height=64, focal=100.0, n_src=2)
    a, b = sc.u8(0), sc.u8(1)
    flat = np.full((64, 64), 128, np.uint8)
    coords = np.concatenate([
        rng.uniform(-4, 68, size=(160, 4)),
        np.repeat(rng.uniform(8, 56, size=(20, 2)), 2, axis=1),   # ref == src position
        rng.uniform(5, 59, size=(20, 4)),
    ])
    out = {}
    for patch in (7, 11):
        r = oracle.compute_ncc("ref", a, b, coords, patch)
        out[f"ncc_valid_{patch}"] = np.array([x.valid for x in r], np.int32)
        out[f"ncc_value_{patch}"] = np.array([x.ncc for x in r])
        out[f"ncc_weak_{patch}"] = np.array([x.weak_texture for x in r], np.int32)
        out[f"ncc_samples_{patch}"] = np.array([x.valid_samples for x in r], np.int32)
    r = oracle.compute_ncc("ref", flat, a, coords[:40], 11)
    out["ncc_flat_valid"] = np.array([x.valid for x in r], np.int32)
    out["ncc_flat_weak"] = np.array([x.weak_texture for x in r], np.int32)
    np.savez(os.path.join(HERE, "ncc.npz"), img_a=a, img_b=b, img_flat=flat, coords=coords, **out)

    # 2. PatchMatchStereo + RefineSubpixel + FuseDepthMaps on a 48x32 scene
    sc = scenes.sphere_arc(width=48, height=32, focal=75.0, n_src=4)
    u8 = np.stack([sc.u8(i) for i in range(5)])
    cams = sc.cams
    res = {}
    for iters, threads, patch in ((0, 1, 7), (1, 1, 7), (2, 3, 5), (2, 1, 7)):
        cfg = N.config_default()
        cfg.patch_size, cfg.iterations, cfg.num_threads = patch, iters, threads
        cfg.min_depth_m, cfg.max_depth_m = 0.25, 0.45
        cfg.random_search_radius_ratio = 0.35
        f, d, s, c = oracle.ref_patchmatch(u8[0], list(u8[1:]), cams[0], cams, cams, [1, 2, 3, 4], cfg)
        key = f"pm_i{iters}_t{threads}_p{patch}"
        res[key + "_cfg"] = cfg_array(cfg)
        res[key + "_filled"] = np.array(f)
        res[key + "_depth"] = d
        res[key + "_support"] = s
        res[key + "_conf"] = c
        if iters == 2 and threads == 1:
            cfg.subpixel_iterations = 5
            n, rd, rc = oracle.ref_refine(u8[0], list(u8[1:]), cams[0], cams, cams, [1, 2, 3, 4],
                                          cfg, d, s, c)
            res["refine_cfg"] = cfg_array(cfg)
            res["refine_count"] = np.array(n)
            res["refine_depth"] = rd
            res["refine_conf"] = rc
            cfg.fusion_voxel_size_m = 0.0005
            fr = oracle.fuse("ref", [(rd, s, rc), (d, s, c)], [cams[0], cams[0]], cfg)
            res["fuse_cfg"] = cfg_array(cfg)
            for k, v in fr.items():
                res["fuse_" + k] = v
    # seeded run: sparse (u, v, depth) points from the truth depth
    cfg = N.config_default()
    cfg.patch_size, cfg.iterations, cfg.num_threads = 5, 1, 2
    cfg.min_depth_m, cfg.max_depth_m = 0.25, 0.45
    ys, xs = np.nonzero(sc.depths[0] > 0)
    pick = rng.choice(len(xs), size=12, replace=False)
    seeds = np.stack([xs[pick] + 0.5, ys[pick] + 0.5, sc.depths[0][ys[pick], xs[pick]]], 1)
    f, d, s, c = oracle.ref_patchmatch(u8[0], list(u8[1:]), cams[0], cams, cams, [1, 2, 3, 4], cfg,
                                       seeds=seeds)
    res.update(seeded_cfg=cfg_array(cfg), seeded_points=seeds, seeded_filled=np.array(f),
               seeded_depth=d, seeded_support=s, seeded_conf=c)
    np.savez(os.path.join(HERE, "patchmatch_48x32.npz"), images=u8, cams=cams_array(cams), **res)

    # 3. SelectSourceViews on the reference's own test poses (test_mvs.cpp:469-500)
    poses = []
    for i in range(5):
        ang = i * 0.15
        c, s = np.cos(ang), np.sin(ang)
        R = np.array([[c, 0, s], [0, 1.0, 0], [-s, 0, c]])
        poses.append(N.Camera.make(400.0, 400.0, 0.0, 0.0, 0, 0, R, [-0.03 * i, 0.0, 0.0]))
    cfg = N.config_default()
    cfg.num_views = 3
    sel = oracle.select_views("ref", poses[0], poses, cfg)
    np.savez(os.path.join(HERE, "select_views.npz"), cams=cams_array(poses), selected=np.array(sel, np.int32))
    print("golden fixtures written")


if __name__ == "__main__":
    main()
