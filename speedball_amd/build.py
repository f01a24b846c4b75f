"""Build libsbgi.so (HIP for gfx950 + host C++) in-tree.

    python -m speedball_amd.build        # or speedball_amd.build.build()

Every .hip source is compiled by hipcc for gfx950 with packed-FP32 instructions
disabled; the host-only scene builder is compiled by g++. Objects go to
speedball_amd/_build/, the library to speedball_amd/libsbgi.so.
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
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libsbgi.so")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SBGI_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 (v_pk_*_f32) is kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE,
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-I", INCLUDE]


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build(verbose: bool = False, jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hip, cpp = _sources()
    cmds = []
    objs = []
    for f in hip:
        o = os.path.join(BUILD, f + ".o")
        objs.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, f), "-o", o])
    for f in cpp:
        o = os.path.join(BUILD, f + ".o")
        objs.append(o)
        cmds.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, f), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    shutil.move(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_addon(verbose)
    build_calib(verbose)
    return LIB


CALIB_DIR = os.path.join(PKG, "calib")


def build_calib(verbose: bool = False):
    """The roofline-ceiling measurement tool (calib/ceilings: VALU issue rates by
    instruction class, L2 and L1 read bandwidth), run by the profiling commands that
    calibrate the walk kernels' rooflines; not part of the library."""
    src = os.path.join(CALIB_DIR, "ceilings.hip")
    exe = os.path.join(CALIB_DIR, "ceilings")
    _run([HIPCC, *HIP_FLAGS, src, "-o", exe + ".tmp"])
    shutil.move(exe + ".tmp", exe)
    if verbose:
        print("built", exe)
    return exe


NAPI_DIR = os.path.join(PKG, "napi")
ADDON = os.path.join(NAPI_DIR, "sbgi.node")
NODE_INCLUDE = os.environ.get("NODE_INCLUDE", "/usr/include/node")


def build_addon(verbose: bool = False):
    """The N-API addon over the C ABI (skipped when the Node headers are absent)."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        if verbose:
            print("node_api.h not found; addon skipped")
        return None
    tmp = ADDON + ".tmp"
    _run(["gcc", "-O2", "-shared", "-fPIC", "-std=c11", "-Wall", "-DNODE_GYP_MODULE_NAME=sbgi", "-I", NODE_INCLUDE,
          os.path.join(NAPI_DIR, "sbgi_napi.c"), "-L", PKG, "-lsbgi", "-Wl,-rpath,$ORIGIN/..", "-o", tmp])
    shutil.move(tmp, ADDON)
    if verbose:
        print("built", ADDON)
    return ADDON


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
