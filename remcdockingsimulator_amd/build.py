"""Build the native library libremc.so (HIP kernels + C ABI) for gfx950.

Usage: python -m remcdockingsimulator_amd.build
The library is written next to this file so that it travels with the source
tree; tests, bench.py and smoke() load it from there and fail loudly when it
is missing.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libremc.so")
SIM = os.path.join(HERE, "remc_sim")
PDBGEN = os.path.join(HERE, "remc_pdbgen")

SOURCES = ["remc_kernels.hip", "remc_analysis.hip", "remc_api.hip"]
FACADE_SOURCES = ["host/remc_facade.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
    "-Wno-unused-command-line-argument",
    "-Wno-unused-value",
]


def hipcc_path():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


# host-side facade and command-line driver: plain C++ over the C ABI
GXX_FLAGS = ["-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unknown-pragmas"]


def build_sim(verbose=False):
    """remc_sim (the REMC driver) and remc_pdbgen (bound-conformation replay),
    linked to libremc.so beside them."""
    inc = ["-I", os.path.join(REPO, "include"), "-I", CSRC, "-I", os.path.join(CSRC, "host")]
    for exe, main in ((SIM, "host/remc_sim.cpp"), (PDBGEN, "host/remc_pdbgen.cpp")):
        srcs = [os.path.join(CSRC, s) for s in (FACADE_SOURCES + [main])]
        tmp = exe + ".tmp"
        cmd = [os.environ.get("CXX", "g++")] + GXX_FLAGS + inc + srcs + ["-L", HERE, "-lremc",
                                                                        "-Wl,-rpath,$ORIGIN", "-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, exe)
    return SIM


def build(verbose=False):
    inc = ["-I", os.path.join(REPO, "include"), "-I", CSRC]
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc_path()] + HIPCC_FLAGS + inc + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_sim(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
