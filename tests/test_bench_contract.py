"""bench.py's contract on the CPU side: --gpus above 1 exits with a message
(this build runs on one GPU), and the argument parser holds the defaults the
bench line reports."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_defaults():
    sys.path.insert(0, ROOT)
    import bench
    old = sys.argv
    try:
        sys.argv = ["bench.py"]
        a = bench.parse()
    finally:
        sys.argv = old
    # the batched mixed-precision workload is the headline; its storage
    # defaults to FP16 inside run_batch
    assert (a.gpus, a.workload, a.images, a.storage, a.width, a.height, a.iterations, a.extras) == \
        (1, "batch", "u8", None, 1920, 1440, 5, 1)


def test_batch_flop_count():
    """The batch roofline's algorithmic FLOPs: P x V x 36 samples x 20 FLOP x
    (1 + 8 x iterations) summed over the jobs, recomputed from its definition
    on a toy job list."""
    sys.path.insert(0, ROOT)
    import bench

    class Cam:
        def __init__(self, w, h):
            self.width, self.height = w, h

    class Sc:
        cams = [Cam(640, 480), Cam(800, 600), Cam(320, 240)]
    jobs = [(0, [1, 2]), (1, [0, 2, 0])]
    want = (640 * 480 * 2 + 800 * 600 * 3) * 36 * 20 * (1 + 8 * 5)
    assert bench.batch_flop(Sc, jobs, 5) == want
