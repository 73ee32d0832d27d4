"""Cooperative cancellation (recon_cancel, recon_c_api.h:601-608; the per
keyframe check of Orchestrator::StageMvs, orchestrator.cpp:989; GPU work is
checked between dispatch batches, docs/01_接口契約.md §3.11).  Every case sets
the flag from the host; every kernel exits normally (the persistent batch
kernel finds its queue exhausted)."""
import ctypes as C
import time

import numpy as np
import pytest

from recon_scanner_amd import _native as N
from recon_scanner_amd import mvs, plane, scenes

pytestmark = pytest.mark.gpu

CANCELLED = 5


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(autouse=True)
def _clean():
    mvs.SetCancelCallback(None)
    mvs.ClearCancel()
    yield
    mvs.SetCancelCallback(None)
    mvs.ClearCancel()


def stage_inputs(n=4, w=160, h=120):
    sc = scenes.sphere_planes(width=w, height=h, focal=2500.0 * w / 1920.0, n_views=n, supersample=1)
    gray = [sc.u8(i) for i in range(n)]
    return sc, gray


def run_stage(sc, gray, mode):
    L, ctx = mvs.lib(), mvs.context()
    n = len(gray)
    gp = (C.c_void_p * n)(*[g.ctypes.data for g in gray])
    cams = (N.Camera * n)(*sc.cams)
    cfg = mvs.MvsConfig(patch_size=7, iterations=2, num_threads=4, min_depth_m=sc.min_depth,
                        max_depth_m=sc.max_depth, num_views=n - 1)
    produced, npts = C.c_int32(-1), C.c_int64(-1)
    st = L.recon_mvs_hip_run_mvs_stage(ctx, n, gp, cams, C.byref(cfg), mode, None, None, None, None,
                                       C.byref(produced), C.byref(npts))
    return st, produced.value


@pytest.mark.parametrize("mode", [0, 1])
def test_stage_cancelled_before_it_starts(mode):
    say("stage cancelled before start, mode", mode)
    sc, gray = stage_inputs()
    mvs.Cancel()
    assert mvs.IsCancelled()
    st, produced = run_stage(sc, gray, mode)
    assert st == CANCELLED and produced == 0
    assert mvs.LastStatus() == CANCELLED
    mvs.ClearCancel()
    assert not mvs.IsCancelled()
    st, produced = run_stage(sc, gray, mode)
    assert st == 0 and produced > 0


def test_callback_cancels_between_keyframes_and_passes():
    """A callback answering yes on its k-th poll stops the reference-mode
    stage at that check point -- before a keyframe or before a directional
    pass -- with status 5; the answer is latched into the flag."""
    sc, gray = stage_inputs()
    polls = []

    def cb():
        polls.append(1)
        return len(polls) >= 6
    mvs.SetCancelCallback(cb)
    st, produced = run_stage(sc, gray, 0)
    say("cancelled after", len(polls), "polls, status", st)
    assert st == CANCELLED and produced == 0
    assert len(polls) == 6
    mvs.SetCancelCallback(None)
    assert mvs.IsCancelled()          # latched
    mvs.ClearCancel()
    st, produced = run_stage(sc, gray, 0)
    assert st == 0 and produced > 0


def test_running_batch_drains_when_cancelled():
    """The persistent batch kernel reads the flag between its work units: a
    batch cancelled shortly after launch finishes far sooner than a full run,
    the next call reports the cancellation, and after clearing a full run
    again equals per-view runs bit for bit."""
    sc, jobs = scenes.orbit_batch(num_refs=24, num_cams=64, seed=0, max_res=(1600, 1200))
    p = plane.default_params(patch_size=11, window_stride=2, top_k=4, iterations=5,
                             min_depth_m=sc.min_depth, max_depth_m=sc.max_depth, value_scale=1.0,
                             tile_schedule=2)
    pm = plane.PlaneMVS(sc.cams, sc.images, storage="f16")
    b = pm.batch(jobs, p)
    b.run()
    pm.synchronize()
    t = time.perf_counter()
    b.run()
    pm.synchronize()
    full = time.perf_counter() - t
    t = time.perf_counter()
    b.run()
    time.sleep(0.25 * full)
    mvs.Cancel()
    pm.synchronize()
    cut = time.perf_counter() - t
    say(f"full run {full * 1e3:.0f} ms, cancelled at {0.25 * full * 1e3:.0f} ms, drained at {cut * 1e3:.0f} ms")
    assert cut < 0.6 * full
    with pytest.raises(RuntimeError):
        b.run()                       # the flag is still set: the call reports it
    assert mvs.LastStatus() == CANCELLED
    mvs.ClearCancel()
    b.run()
    pm.synchronize()
    r, srcs = jobs[0]
    g = pm.depth_map(r)
    pm.patchmatch(r, srcs, p)
    s = pm.depth_map(r)
    for k in ("depth", "normals", "cost", "support"):
        assert g[k].tobytes() == s[k].tobytes(), k
    b.close()
