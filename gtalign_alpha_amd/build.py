"""In-tree build of the native libraries.

libgta_amd.so    HIP kernels + C ABI (hipcc, gfx950)
libgta_synth.so  synthetic-workload generator (host C++)
The oracle (test infrastructure) is built by oracle/Makefile.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgta_amd.so")
SYNTH = os.path.join(HERE, "libgta_synth.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GTA_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    # packed FP32 instructions stay off (the compiler would otherwise form them)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC,
]
HIP_SOURCES = ["gta_kernels.hip", "gta_capi.hip"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip(out=LIB):
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + [os.path.join(CSRC, s) for s in HIP_SOURCES] + ["-o", tmp])
    os.replace(tmp, out)


def build_synth(out=SYNTH):
    cxx = shutil.which("g++") or "g++"
    tmp = out + ".tmp"
    _run([cxx, "-std=c++17", "-O2", "-fPIC", "-shared", "-fopenmp", os.path.join(CSRC, "synth.cpp"), "-o", tmp])
    os.replace(tmp, out)


def build_oracle(force=False):
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8"] + (["-B"] if force else []) + ["all"])


def build_all():
    build_synth()
    build_oracle()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "synth": build_synth, "oracle": build_oracle}[w]()
