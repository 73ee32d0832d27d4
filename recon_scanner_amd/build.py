"""Builds the native libraries of the package in-tree.

  _lib/librecon_mvs_hip.so  HIP kernels for gfx950 + the C ABI (include/recon_mvs_hip.h)
                            + the C++ recon:: entry points (include/recon/mvs.h)
  _lib/libreconscene.so     synthetic scene renderer (host C++)

Every HIP file is compiled with packed-FP32 instructions disabled.  Files that
must reproduce the reference's double arithmetic bit for bit are compiled with
-ffp-contract=off.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "_lib")
OBJ = os.path.join(ROOT, "_build", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYTORCH_ROCM_ARCH", "gfx950")

HIP_FLAGS = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-I", INC, "-I", CSRC]
# (source, extra flags)
HIP_SOURCES = [
    ("context.hip", []),
    ("ref_mode.hip", ["-ffp-contract=off"]),
    ("fuse.hip", ["-ffp-contract=off"]),
    ("plane_mode.hip", ["-ffp-contract=off"]),
    ("plane_adjoint.hip", ["-ffp-contract=off"]),
    ("mvs_api.cpp", ["-ffp-contract=off"]),
    ("mvs_stage.cpp", ["-ffp-contract=off"]),
    ("spill.cpp", []),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return r


def _compile(src, extra, obj_dir=OBJ):
    out = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + HIP_FLAGS + extra
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", out]
    _run(cmd)
    return out


def build_hip(jobs=8, defines=(), out_dir=LIB, obj_dir=OBJ):
    """Compiles every HIP source; `defines` (-D...) are added to each compile."""
    os.makedirs(obj_dir, exist_ok=True)
    os.makedirs(out_dir, exist_ok=True)
    srcs = [(s, list(e) + list(defines)) for s, e in HIP_SOURCES
            if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda se: _compile(se[0], se[1], obj_dir), srcs))
    tmp = os.path.join(out_dir, "librecon_mvs_hip.so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, os.path.join(out_dir, "librecon_mvs_hip.so"))


def build_scene():
    os.makedirs(LIB, exist_ok=True)
    tmp = os.path.join(LIB, "libreconscene.so.tmp")
    _run(["g++", "-std=c++17", "-O2", "-fopenmp", "-fPIC", "-shared", "-I", INC,
          os.path.join(CSRC, "scene.cpp"), "-o", tmp])
    os.replace(tmp, os.path.join(LIB, "libreconscene.so"))


def build_cpp_test():
    """tests/cpp/test_mvs_api.cpp: a recon:: API caller linked against the library."""
    out_dir = os.path.join(ROOT, "_build")
    os.makedirs(out_dir, exist_ok=True)
    tmp = os.path.join(out_dir, "test_mvs_api.tmp")
    _run(["g++", "-std=c++17", "-O2", "-I", INC, os.path.join(ROOT, "tests", "cpp", "test_mvs_api.cpp"),
          "-L", LIB, "-lrecon_mvs_hip", "-lreconscene", f"-Wl,-rpath,{LIB}",
          "-Wl,-rpath,$ORIGIN/../recon_scanner_amd/_lib", "-o", tmp])
    os.replace(tmp, os.path.join(out_dir, "test_mvs_api"))


def build_all():
    build_scene()
    build_hip()
    build_cpp_test()


if __name__ == "__main__":
    build_all()
