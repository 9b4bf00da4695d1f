"""Builds the native libraries of jagua-rs_amd in-tree.

    python jagua-rs_amd/build.py            # library + synthetic-workload generator
    python jagua-rs_amd/build.py --oracle   # also the CPU oracle (test infrastructure)

Outputs:
    jagua-rs_amd/lib/libjagua_amd.so    C ABI of include/jagua_amd.h (host build + gfx950 kernels)
    jagua-rs_amd/lib/libjagua_synth.so  synthetic nesting workloads (host only)
    oracle/_build/liboracle.so          CPU oracle, built by oracle/Makefile

Device code is compiled for gfx950 only, with -ffp-contract=off (no FMA contraction: the
reference never fuses), IEEE f32 division/sqrt (the HIP default), f32 denormals kept
(-fno-gpu-flush-denormals-to-zero, also the default, stated so that no toolchain default can
change it), and packed-FP32 instruction selection disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "lib", "obj")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-fno-builtin-sinf", "-fno-builtin-cosf", "-fno-builtin-sincosf",
    "-Wall", "-Wno-unused-function", "-Wno-unused-parameter",
]
DEVICE = [
    f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-flush-denormals-to-zero", "-munsafe-fp-atomics",
]

SOURCES = ["jg_host.cpp", "jg_api.hip", "jg_batched.hip", "jg_overlap.hip", "jg_poles.hip", "jg_search.hip", "jg_piers.hip", "jg_collect.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_library(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cmds = []
    objs = []
    for s in srcs:
        o = os.path.join(OBJ, s + ".o")
        objs.append(o)
        cmds.append([HIPCC, "-x", "hip", *COMMON, *DEVICE, "-c", os.path.join(CSRC, s), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(LIB, "libjagua_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, *COMMON, *DEVICE, "-shared", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_synth():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libjagua_synth.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
          "-fno-builtin-sinf", "-fno-builtin-cosf", "-fno-builtin-sincosf",
          os.path.join(CSRC, "jg_synth.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle():
    if shutil.which("make") is None:
        raise RuntimeError("make not found")
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B"])
    return os.path.join(ROOT, "oracle", "_build", "liboracle.so")


def main(argv):
    outs = [build_synth(), build_library()]
    if "--oracle" in argv:
        outs.append(build_oracle())
    for o in outs:
        print("built", os.path.relpath(o, ROOT))


if __name__ == "__main__":
    main(sys.argv[1:])
