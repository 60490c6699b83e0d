"""Build the gfx950 engine library and CLI in-tree.

    python dmrg-implementations_amd/build.py [--jobs N]

Produces
  dmrg-implementations_amd/dmrg_amd/libdmrg_amd.so   (C-ABI of include/dmrg_amd.h)
  dmrg-implementations_amd/bin/dmrg2_gpu_amd         (reference-compatible CLI)
Objects go to dmrg-implementations_amd/_build/ (not committed).
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "dmrg_amd", "libdmrg_amd.so")
BIN = os.path.join(HERE, "bin", "dmrg2_gpu_amd")

SOURCES = ["gemm.hip", "contract.hip", "lanczos.hip", "svd.hip", "warm.hip", "engine.hip", "pdmrg.hip", "noalloc.hip",
           "capi.hip"]
CLI = ["cli.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed-FP32 VALU ops are disabled for every HIP source.  MFMA accumulators
# are kept in VGPRs (VGPR form of the MFMA instructions): with AGPR
# accumulators the compiler copies every accumulator out of and back into the
# AGPRs at each iteration of the GEMM k-loop; measured +5-8 % on the FP64
# H_eff GEMMs, neutral on the FP32 ones.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-amdgpu-mfma-vgpr-form"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        out = _run([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    cli_src = [os.path.join(CSRC, s) for s in CLI if os.path.exists(os.path.join(CSRC, s))]
    if cli_src:
        _run([HIPCC] + FLAGS + cli_src + ["-o", BIN, "-L" + os.path.dirname(LIB), "-ldmrg_amd",
                                          "-Wl,-rpath,$ORIGIN/../dmrg_amd"])
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
    sys.exit(0)
