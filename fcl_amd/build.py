"""Builds libfclx.so (HIP kernels for gfx950 + host C++) in-tree.

Every .hip file is compiled by hipcc for gfx950, host-only .cpp files by the
same driver as plain C++, all with -ffp-contract=off so that no multiply-add
is fused (the FP64 paths are bit-exact only without contraction) and with
packed-FP32 instructions disabled.  The shared library lands next to this
file, where fcl_amd._native loads it.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfclx.so")

HIP_SOURCES = ["api.hip", "kernels_plain.hip", "kernels_front.hip", "kernels_refit.hip",
               "kernels_broadphase.hip",
               "kernels_shape.hip",
               "kernels_meshshape.hip",
               "engine.hip", "engine_dist.hip", "host/fcl_api.hip", "host/serve.hip"]
HOST_SOURCES = ["host/bvh_build.cpp", "host/synth.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-I", os.path.join(ROOT, "include")]
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP engine cannot be built")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_library(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
        cmds.append([hipcc] + COMMON + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
        cmds.append(["g++"] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if out.strip():
                sys.stderr.write(out + "\n")
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library())
