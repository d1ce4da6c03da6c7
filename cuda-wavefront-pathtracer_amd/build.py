"""Builds libcwp.so (C++17 host code + gfx950 HIP kernels), the `pt` CLI and the headless `viewer` in-tree.

Host sources compile with g++, kernels with hipcc for gfx950 only. Both use -ffp-contract=off: fused multiply-adds
are written explicitly in the sources, so host and device evaluate the same floating-point operations. Packed FP32
instructions are disabled for device code.

    python cuda-wavefront-pathtracer_amd/build.py [--jobs N] [--quiet] [--incremental]
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libcwp.so")
CLI = os.path.join(PKG, "pt")
VIEWER = os.path.join(PKG, "viewer")

HOST_SOURCES = ["host/mesh.cpp", "host/camera.cpp", "host/io.cpp", "host/texture.cpp", "host/bvh.cpp", "host/scene_check.cpp", "host/instances.cpp", "host/envmap.cpp", "host/bvhcache.cpp", "host/setup.cpp", "host/serve.cpp", "capi.cpp"]
HIP_SOURCES = ["device/renderer.hip", "device/common.hip", "device/kernels_v2.hip", "device/kernels_v4.hip", "device/kernels_v8.hip",
               "device/hploc.hip", "device/collapse.hip", "device/batch.hip", "device/display.hip"]
CLI_SOURCES = ["app/pt.cpp", "app/viewer.cpp"]

INCLUDES = ["-I" + CSRC, "-I" + os.path.join(ROOT, "include")]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libcwp.so")


def _run(cmd, quiet):
    if not quiet:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed (%d):\n%s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    if r.stdout.strip() and not quiet:
        print(r.stdout, flush=True)


def _headers_mtime():
    t = 0.0
    for d in (CSRC, os.path.join(ROOT, "include")):
        for base, _, files in os.walk(d):
            for f in files:
                if f.endswith(".h"):
                    t = max(t, os.path.getmtime(os.path.join(base, f)))
    return t


def _stale(src, obj, headers):
    """incremental builds: an object is rebuilt when it is older than its source or than any header"""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return t < os.path.getmtime(src) or t < headers


def build(jobs=8, quiet=False, incremental=False):
    os.makedirs(BUILD, exist_ok=True)
    hc = hipcc()
    steps = []
    objs = []
    headers = _headers_mtime() if incremental else 0.0
    for s in HOST_SOURCES + CLI_SOURCES:
        o = os.path.join(BUILD, s.replace("/", "_") + ".o")
        if not incremental or _stale(os.path.join(CSRC, s), o, headers):
            steps.append(["g++"] + HOST_FLAGS + INCLUDES + ["-c", os.path.join(CSRC, s), "-o", o])
        if s not in CLI_SOURCES:
            objs.append(o)
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s.replace("/", "_") + ".o")
        if not incremental or _stale(os.path.join(CSRC, s), o, headers):
            steps.append([hc] + HIP_FLAGS + INCLUDES + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for f in [ex.submit(_run, c, quiet) for c in steps]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hc, "-shared", "-fPIC", "-o", tmp] + objs, quiet)
    os.replace(tmp, LIB)
    for src, exe in (("app/pt.cpp", CLI), ("app/viewer.cpp", VIEWER)):
        cli_obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
        _run([hc, "-o", exe + ".tmp", cli_obj, "-L" + PKG, "-lcwp", "-Wl,-rpath,$ORIGIN"], quiet)
        os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--incremental", action="store_true", help="rebuild only objects older than their source or a header")
    a = ap.parse_args()
    try:
        print(build(a.jobs, a.quiet, a.incremental))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
