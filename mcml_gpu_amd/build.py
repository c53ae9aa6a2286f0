"""Build libmcml.so (HIP kernels + C ABI + I/O) and the MCML_GPU executable for gfx950.

Everything is compiled in-tree with hipcc; no JIT cache is involved, so the built files
travel with the source tree.  Usage: python -m mcml_gpu_amd.build
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmcml.so")
EXE = os.path.join(HERE, "bin", "MCML_GPU")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off and no fast-math: the transport must round exactly like its CPU
# restatement.  -packed-fp32-ops: no v_pk_* FP32 instructions are generated.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-Wno-unused-result", "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = ["transport.hip", "sort.hip", "white.hip", "mlp.hip", "runtime.cpp", "mlp_host.cpp", "io.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(EXE), exist_ok=True)

    def obj(src):
        o = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
        extra = DEVICE + (["-x", "hip"] if src.endswith(".hip") else [])
        _run([HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", o])
        return o

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(obj, SOURCES))
    _run([HIPCC] + DEVICE + ["-shared", "-fPIC", "-o", LIB + ".tmp"] + objs)
    os.replace(LIB + ".tmp", LIB)
    _run([HIPCC, "-O2", "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "cli_main.cpp"),
          "-o", EXE, "-L" + HERE, "-lmcml", "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print("built", LIB, "and", EXE)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
