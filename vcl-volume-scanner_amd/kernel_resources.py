"""Register, LDS and scratch use of every kernel of the fusion library, read
from the gfx950 code objects' own metadata (the AMDGPU ELF notes the compiler
writes), with the waves per SIMD those allow.

    python vcl-volume-scanner_amd/kernel_resources.py [--md]

Compiles each HIP source for the device only, with the library's own flags
(build.py), into _build/, and prints one row per kernel. The VGPR column is
`.vgpr_count` (architectural VGPRs, allocated in blocks of 8 of the 512 per
SIMD lane); a profiler's per-dispatch VGPR figure can use another unit.
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import build  # noqa: E402

LLVM = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "llvm", "bin")
KEYS = ("vgpr_count", "agpr_count", "sgpr_count", "vgpr_spill_count", "sgpr_spill_count",
        "group_segment_fixed_size", "private_segment_fixed_size", "max_flat_workgroup_size")


def code_object(src):
    os.makedirs(build.BUILD, exist_ok=True)
    base = os.path.join(build.BUILD, src.replace(".hip", ""))
    bundle, co = base + ".dev.o", base + f".{build.ARCH}.co"
    build._run([build.HIPCC, *build.hip_flags(src), "--offload-device-only", "-c", os.path.join(build.CSRC, src),
                "-o", bundle])
    build._run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", f"--input={bundle}",
                f"--targets=hipv4-amdgcn-amd-amdhsa--{build.ARCH}", f"--output={co}"])
    return co


def kernels(co):
    notes = build._run([os.path.join(LLVM, "llvm-readelf"), "--notes", co]).stdout
    out, cur = [], None
    for line in notes.splitlines():
        m = re.match(r"\s*-?\s*\.(\w+):\s+(\S+)\s*$", line)
        if not m:
            continue
        k, v = m.group(1), m.group(2)
        if k == "agpr_count" and line.lstrip().startswith("-"):
            cur = {}
            out.append(cur)
        if cur is None:
            continue
        if k == "name":
            cur["name"] = v
        elif k in KEYS:
            cur[k] = int(v)
    return [k for k in out if "name" in k and not k["name"].endswith(".kd")]


def demangle(names):
    r = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True)
    return r.stdout.splitlines()


def waves_per_simd(k):
    regs = k.get("vgpr_count", 0) + k.get("agpr_count", 0)
    by_vgpr = 512 // max(8, -(-regs // 8) * 8)
    lds = k.get("group_segment_fixed_size", 0)
    wg = max(1, k.get("max_flat_workgroup_size", 256))
    waves_per_wg = -(-wg // 64)
    by_lds = 8 if lds == 0 else (160 * 1024 // lds) * waves_per_wg // 4
    return min(8, by_vgpr, max(1, by_lds)), by_vgpr


def main():
    md = "--md" in sys.argv
    rows = []
    for src in build.HIP_SOURCES:
        ks = kernels(code_object(src))
        for k, name in zip(ks, demangle([k["name"] for k in ks])):
            occ, by_vgpr = waves_per_simd(k)
            rows.append((src, name.replace("vs::", ""), k.get("vgpr_count", 0), k.get("agpr_count", 0),
                         k.get("sgpr_count", 0), k.get("vgpr_spill_count", 0), k.get("sgpr_spill_count", 0),
                         k.get("group_segment_fixed_size", 0), k.get("private_segment_fixed_size", 0),
                         min(8, by_vgpr), occ))
    head = ("source", "kernel", "VGPR", "AGPR", "SGPR", "VGPR spills (to memory)", "SGPR spills (to VGPR lanes)",
            "LDS B", "scratch B", "waves/SIMD (VGPR)", "waves/SIMD (VGPR+LDS)")
    if md:
        print("| " + " | ".join(head) + " |")
        print("|" + "---|" * len(head))
        for r in rows:
            print("| " + " | ".join(str(x) for x in r) + " |")
    else:
        for r in rows:
            print("\t".join(str(x) for x in r))


if __name__ == "__main__":
    main()
