"""Compile the MI355X MSM library and the msmpb-compatible command-line tool.

Outputs (in-tree, git-ignored):
  msmpb_amd/libmsmpb_hip.so  -- HIP kernels + forcefield.h C ABI (hipcc, gfx950)
  msmpb_amd/libpme_hip.so    -- the PME comparator's pme.h C ABI on those kernels
  msmpb_amd/bin/msmpb        -- CLI with the reference msmpb's argv and JSON
  msmpb_amd/bin/pme          -- CLI with the reference pme's argv and JSON
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmsmpb_hip.so")
CLI = os.path.join(HERE, "bin", "msmpb")
PME_LIB = os.path.join(HERE, "libpme_hip.so")
PME_CLI = os.path.join(HERE, "bin", "pme")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MSMPB_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["msm_engine.hip", "msm_kernels.hip", "msm_mfma.hip", "msm_stencil.hip",
               "msm_plan.cpp"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 arithmetic stays disabled on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
    # calls between the library's own entry points stay inside it, also when
    # libpme_hip.so (same FF_* names, other signatures) is loaded first
    "-Wl,-Bsymbolic-functions",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build_library(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *srcs], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_cli(verbose=False):
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    _run(["gcc", "-O2", "-std=c11", "-Wall", "-I", os.path.join(ROOT, "include"),
          "-o", CLI, os.path.join(CSRC, "msmpb_cli.c"), "-L", HERE, "-lmsmpb_hip",
          "-Wl,-rpath,$ORIGIN/..", "-lm"], verbose)
    return CLI


def build_pme(verbose=False):
    inc = ["-I", os.path.join(ROOT, "include")]
    tmp = PME_LIB + ".tmp"
    _run(["gcc", "-O2", "-std=gnu11", "-Wall", "-fPIC", "-shared", *inc, "-o", tmp,
          os.path.join(CSRC, "pme_capi.c"), "-L", HERE, "-lmsmpb_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-Bsymbolic-functions",
          "-lm"], verbose)
    os.replace(tmp, PME_LIB)
    _run(["gcc", "-O2", "-std=c11", "-Wall", *inc, "-o", PME_CLI, os.path.join(CSRC, "pme_cli.c"),
          "-L", HERE, "-lpme_hip", "-lmsmpb_hip", "-Wl,-rpath,$ORIGIN/..", "-lm"], verbose)
    return PME_LIB


def build(verbose=False):
    build_library(verbose)
    build_cli(verbose)
    build_pme(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
    print(CLI)
    print(PME_LIB)
    print(PME_CLI)
