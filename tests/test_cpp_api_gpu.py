"""The C++ drop-in: a program written against include/recon/mvs.h (the
reference's interface) and linked to librecon_mvs_hip.so runs the reference's
unit-test scene with the reference's thresholds."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "_build", "test_mvs_api")

pytestmark = pytest.mark.gpu


def test_cpp_api_caller_passes():
    assert os.path.exists(BIN), "run __graft_entry__.build() first"
    print("[case] _build/test_mvs_api", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
