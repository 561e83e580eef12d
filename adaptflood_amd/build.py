"""Build recipe for the native library (gfx950 HIP + C++ host code).

Produces, in-tree:
  adaptflood_amd/_lib/libadaptflood_amd.so   the C ABI of include/adaptflood.h
  adaptflood_amd/_lib/lisflood_amd           the command-line driver
  adaptflood_amd/_lib/af_hbm_probe           the bandwidth probe of the step's memory pattern

Every HIP translation unit is compiled with -ffp-contract=off (the FP64 path
must evaluate the reference's operations without contracted FMAs) and with
packed-FP32 instruction selection disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AF_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-pthread"]

HIP_SOURCES = ["compat.hip", "fused.hip"]
CXX_SOURCES = ["io.cpp", "common.cpp", "ncwriter.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} {os.path.basename(cmd[-1])}")
    return r.stdout


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run(["g++", *CXX_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    lib = os.path.join(OUT, "libadaptflood_amd.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp, "-lpthread"])
    os.replace(tmp, lib)
    drv_src = os.path.join(CSRC, "driver.cpp")
    if os.path.exists(drv_src):
        drv_obj = os.path.join(BUILD, "driver.cpp.o")
        _run(["g++", *CXX_FLAGS, *inc, "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
              "-c", drv_src, "-o", drv_obj])
        exe = os.path.join(OUT, "lisflood_amd")
        _run(["g++", drv_obj, "-o", exe, "-L", OUT, "-ladaptflood_amd",
              "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    # the bandwidth probe of the step's memory pattern (bench.py --workload ceiling)
    probe_src = os.path.join(CSRC, "hbm_probe.hip")
    if os.path.exists(probe_src):
        _run([HIPCC, *HIP_FLAGS, probe_src, "-o", os.path.join(OUT, "af_hbm_probe")])
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build(verbose=True)
