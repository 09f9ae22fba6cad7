"""Builds libmhd2d.so (the HIP library behind include/mhd2d.h) for gfx950.

Run as ``python -m resistivemhd_gpu_amd.build`` or call ``build()``.  The
library lands next to this file so that it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmhd2d.so")

SOURCES = ["mhd2d_kernels.hip", "mhd2d_capi.hip", "mhd2d_batch.hip", "mhd2d_adjoint.hip", "mhd2d_setups.cpp"]

# -ffp-contract=off: no fused multiply-add contraction, so each operation
# rounds as written in the reference's expressions.
# -packed-fp32-ops off: no packed FP32 instructions anywhere in the library.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, out=LIB, defines=()):
    """Compiles the library to `out`; `defines` are extra -D settings
    (e.g. MHD_WAVES_PER_EU_EXACT=1) for launch-shape experiments."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-D" + d for d in defines] + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libmhd2d.so")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
