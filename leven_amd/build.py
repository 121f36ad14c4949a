"""Build leven_amd/liblvn_mesher.so for gfx950 with hipcc (cross-compiles without a GPU).

    python -m leven_amd.build          # compile every translation unit, link the library
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblvn_mesher.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["lvn_kernels.hip", "lvn_runtime.hip", "lvn_server.hip", "lvn_seams.hip", "lvn_clipmap.cpp",
           "lvn_simplify.hip"]

# -ffp-contract=off: the kernels follow the reference's operation order with fused
# multiply-adds only where written as fmaf(), as the IEEE FP32 restatement does,
# so results are bit-identical to it.
# Packed FP32 instructions are disabled on this platform.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # the host pass does not know the packed-fp32 target feature: drop that warning
    noise = [ln for ln in res.stdout.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    return src, res.returncode, "\n".join(noise), out


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    failed = False
    for src, rc, log, _ in results:
        if log and verbose:
            print(f"[{src}]\n{log}", file=sys.stderr)
        failed = failed or rc != 0
    if failed:
        raise RuntimeError("hipcc compilation failed")
    objs = [o for *_, o in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
