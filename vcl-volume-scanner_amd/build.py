"""Build recipe for the fusion library (in-tree, for gfx950).

    python vcl-volume-scanner_amd/build.py

produces, next to this file,
  libvscan.so        HIP kernels + the C ABI of include/vscan.h (hipcc, gfx950)
  libvscan_scene.so  the synthetic depth-stream generator of include/vscan_scene.h (g++)
Objects go to _build/. Every HIP source is compiled with -ffp-contract=off
(index and mask arithmetic must round once per operation) and with packed
FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VS_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["vs_img.hip", "vs_icp.hip", "vs_vol.hip", "vs_pre.hip", "vs_api.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result", "-Wno-unused-value",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

# Instruction-scheduling strategy of the AMDGPU backend per source (the
# default aims for occupancy): the ICP's latency-bound iteration gains from
# the ILP-maximising schedule (ICP stage 1.98 -> 1.87 ms per step), the
# raycast shading from the iterative one (0.63 -> 0.60 ms); the march (same
# file as the shading) is unchanged within noise, and the bilateral filter
# loses with either (0.19 -> 0.21-0.27 ms), so vs_img.hip keeps the default.
FILE_FLAGS = {
    "vs_icp.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
    "vs_vol.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"],
}


def hip_flags(src):
    return HIP_FLAGS + FILE_FLAGS.get(src, [])


LIB = os.path.join(HERE, "libvscan.so")
SCENE_LIB = os.path.join(HERE, "libvscan_scene.so")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


# Variant builds for measured comparisons (`python build.py --variant NAME`
# writes libvscan_NAME.so next to libvscan.so; tests and the benchmark load
# libvscan.so, or the file VSCAN_LIB names):
#   icp_lds: the ICP iteration's block sums as per-thread FP32 FMAs and LDS
#            chunk sums instead of on the matrix cores (VS_ICP_MFMA=0)
VARIANTS = {"icp_lds": ["-DVS_ICP_MFMA=0"]}


def build(verbose=False, variant=None):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    extra = VARIANTS[variant] if variant else []
    tag = f"_{variant}" if variant else ""
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", f"{tag}.o"))
        objs.append(obj)
        cmds.append([HIPCC, *hip_flags(src), *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        for r in ex.map(_run, cmds):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    lib = LIB.replace(".so", f"{tag}.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, lib)
    if variant:
        return lib, None
    tmp = SCENE_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", f"-I{INCLUDE}",
          os.path.join(CSRC, "scene.cpp"), "-o", tmp])
    os.replace(tmp, SCENE_LIB)
    return LIB, SCENE_LIB


if __name__ == "__main__":
    var = sys.argv[sys.argv.index("--variant") + 1] if "--variant" in sys.argv else None
    print(build(verbose="-v" in sys.argv, variant=var))
