"""Build libkpm_amd.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m kwant_amd.build  (or kwant_amd.build.build()).
Each translation unit is compiled on its own (in parallel) into a build
folder next to this file, then linked; the library is written next to this
file so it travels with the source tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkpm_amd.so")
SOURCES = ["kpm_capi.hip", "kpm_cheb.hip", "kpm_gemm.hip", "kpm_mlp.hip", "kpm_shift.hip",
           "kpm_lattice.hip", "kpm_sliced.hip", "kpm_static.hip", "kpm_tiled.hip",
           "kpm_march.hip"]
ARCH = os.environ.get("KPM_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-munsafe-fp-atomics",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
    return res.returncode


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    objs = [os.path.join(OBJ, s.replace(".hip", ".o")) for s in SOURCES]
    cmds = [[HIPCC] + FLAGS + ["-c", os.path.join(CSRC, s), "-o", o]
            for s, o in zip(SOURCES, objs)]
    jobs = jobs or min(len(cmds), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        codes = list(pool.map(lambda c: _run(c, verbose), cmds))
    if any(codes):
        raise RuntimeError("hipcc failed compiling libkpm_amd sources")
    tmp = LIB + ".tmp"
    if _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", tmp],
            verbose):
        raise RuntimeError("hipcc failed linking libkpm_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
