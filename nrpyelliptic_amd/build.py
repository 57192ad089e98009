"""Build libnrpyell_hip.so in-tree with hipcc for gfx950.

    python -m nrpyelliptic_amd.build          # or __graft_entry__.build()
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnrpyell_hip.so")
SOURCES = ["nrpyell_host.cpp", "nrpyell_kernels.hip", "nrpyell_substep.hip", "nrpyell_substep2.hip",
           "nrpyell_capi.hip", "nrpyell_mlp.hip", "nrpyell_adm.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # every FMA in the kernels is written out; no implicit contraction
    "-ffp-contract=off",
    # keep the compiler from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # register-pressure tracking of the AMDGPU scheduler: the stages-3+4 two-substep kernel at
    # 168 VGPRs spills 6 instead of 10 and runs 9.5 instead of 9.8 us at 128x128x16
    "-mllvm", "-amdgpu-use-amdgpu-trackers",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wl,-Bsymbolic",
    "-Wl,--no-undefined",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(HERE, "..", "include"), "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libnrpyell_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
