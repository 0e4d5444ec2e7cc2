"""Build of the native parts: libaero3d.so (HIP, gfx950) and the Node-API addon.

    python 3d_aero_lab_amd/build.py          # build everything in-tree

Outputs land beside the sources (3d_aero_lab_amd/lib/), so they travel with
the tree to the GPU host; they are git-ignored.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled in every HIP object.
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result", "-munsafe-fp-atomics"]
LIBAERO = os.path.join(LIB, "libaero3d.so")
ADDON = os.path.join(LIB, "aero3d.node")


def _run(cmd):
    r = subprocess.run(cmd, cwd=PKG, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {cmd[0]} exit {r.returncode}")
    return r


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    src = [os.path.join(CSRC, f) for f in ("aero3d.hip", "aero3d_batch.hip", "axi.hip", "field.hip")]
    tmp = LIBAERO + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *src])
    os.replace(tmp, LIBAERO)
    return LIBAERO


def build_addon():
    """Node-API addon: plain C++ against the Node headers, linked to libaero3d.so."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")):
        return None
    tmp = ADDON + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-DNODE_GYP_MODULE_NAME=aero3d", "-I", inc, "-I", os.path.join(ROOT, "include"),
          "-o", tmp, os.path.join(CSRC, "aero3d_napi.cc"), "-L", LIB, "-laero3d", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, ADDON)
    return ADDON


def build_all():
    out = [build_hip()]
    if os.path.exists(os.path.join(CSRC, "aero3d_napi.cc")):
        out.append(build_addon())
    return out


if __name__ == "__main__":
    for p in build_all():
        print(p)
