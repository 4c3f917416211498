"""Compile liblmhip.so for gfx950 with hipcc.

The library is built in-tree (next to this package) so that it travels with the source tree.
Flags:
  -ffp-contract=off   no multiply-add is fused that the reference (Julia) rounds twice; the
                      per-particle parity with the CPU kernels depends on it
  -packed-fp32-ops    packed FP32 instructions are kept out of the device code
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))   # langevinmedium.jl_amd/
REPO_ROOT = os.path.dirname(PKG_ROOT)
CSRC = os.path.join(PKG_ROOT, "csrc")
LIB_PATH = os.path.join(PKG_ROOT, "liblmhip.so")
SOURCES = [os.path.join(CSRC, "lmhip.hip")]
DEPS = SOURCES + [os.path.join(CSRC, f) for f in ("lmhip_device.h", "lmhip_adjoint.h", "lmhip_batch.h", "lmhip_sample.h")] + [
    os.path.join(REPO_ROOT, "include", "lmhip.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (expected /opt/rocm/bin/hipcc)")


def build(verbose: bool = False) -> str:
    """Compile the HIP library; always recompiles.  Returns the path of the .so."""
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def is_stale() -> bool:
    if not os.path.exists(LIB_PATH):
        return True
    t = os.path.getmtime(LIB_PATH)
    return any(os.path.getmtime(d) > t for d in DEPS)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
