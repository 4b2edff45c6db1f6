"""Build recipe of the native libraries (hipcc for gfx950, g++ for host code).

  libmvdecon_hip.so  -- the product: HIP kernels + the C ABI of include/mvdecon.h
  libmvd_synth.so    -- host-only seeded synthetic data generator

Run as ``python -m spim_registration_amd.build`` or through
``__graft_entry__.build()``.  Objects go to ``_build/`` (git-ignored), the
libraries next to this file.
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
BUILD = os.path.join(ROOT, "_build")

HIP_SOURCES = ["conv_kernels.hip", "fft_fast.hip", "fft_xy.hip", "pointwise_kernels.hip", "gradient.hip", "session.hip", "jna_abi.hip", "transform.hip", "psf.hip", "sepconv.hip", "export.hip", "staging.hip"]
HIP_LIB = os.path.join(PKG, "libmvdecon_hip.so")
SYNTH_LIB = os.path.join(PKG, "libmvd_synth.so")

ARCH = os.environ.get("MVD_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout)
    # hipcc's host pass warns about the GPU-only target feature; keep other output
    lines = [ln for ln in r.stdout.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")


def build_hip(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_synth() -> str:
    tmp = SYNTH_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall",
          os.path.join(CSRC, "synth.cpp"), "-o", tmp])
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


def build_all() -> None:
    build_synth()
    build_hip()


if __name__ == "__main__":
    build_all()
    print(HIP_LIB)
    print(SYNTH_LIB)
