"""Build recipe for the in-tree native libraries.

    python -m merge_fusion_amd.build            # everything
    python -m merge_fusion_amd.build --no-oracle

Produces, next to their sources:
    merge_fusion_amd/libkifi_hip.so    HIP kernels + C ABI (hipcc, gfx950)
    merge_fusion_amd/libkifi_synth.so  seeded procedural depth frames (gcc)
    oracle/liboracle.so                CPU oracle (test infrastructure, gcc)

The HIP code is compiled with -ffp-contract=off so binary32 arithmetic follows
source order (the parity contract with the oracle) and with packed-FP32
instruction selection disabled.
"""
import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")

HIP_SOURCES = ["kifi_api.hip", "kifi_integrate.hip", "kifi_mesh.hip", "kifi_icp.hip", "kifi_pipeline.hip",
               "kifi_learn.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", INCLUDE, "-I", CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build_hip(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in srcs]

    def one(pair):
        src, obj = pair
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(one, zip(srcs, objs)))
    out = os.path.join(PKG, "libkifi_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_synth():
    out = os.path.join(PKG, "libkifi_synth.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-std=c11", "-I", INCLUDE,
          os.path.join(CSRC, "synth.c"), "-o", out, "-lm"])
    return out


def build_cpp_tests():
    """The C++ drop-in test program (host code only: g++ against the ABI)."""
    src = os.path.join(PKG, "cpp", "tests", "test_dropin.cpp")
    out = os.path.join(PKG, "cpp", "tests", "test_dropin")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wno-unused-result", "-I", os.path.join(PKG, "cpp"), "-I", INCLUDE,
          src, "-o", out, "-L", PKG, "-lkifi_hip", "-lkifi_synth", "-Wl,-rpath,$ORIGIN/../..",
          "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"])
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_all(oracle=True):
    outs = [build_synth(), build_hip(), build_cpp_tests()]
    if oracle:
        outs.append(build_oracle())
    return outs


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args(argv)
    for o in build_all(oracle=not a.no_oracle):
        print("built", os.path.relpath(o, ROOT))


if __name__ == "__main__":
    main()
