"""Build the CPU oracle (liblmoracle.so) with gcc.

TEST INFRASTRUCTURE ONLY: the oracle is the checker the HIP backend is compared against.
-ffp-contract=off keeps every product and sum rounded separately, as Julia does.

The reference itself (Julia) cannot be built or run here, so there is no oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblmoracle.so")
SOURCES = [os.path.join(HERE, "lm_oracle.c")]
DEPS = SOURCES + [os.path.join(HERE, "lm_oracle.h"),
                  os.path.join(os.path.dirname(HERE), "include", "lmhip.h")]
CFLAGS = ["-O2", "-fPIC", "-shared", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          "-Wall", "-Wextra"]


def build(verbose: bool = False) -> str:
    tmp = LIB_PATH + ".tmp"
    cmd = [os.environ.get("CC", "gcc"), *CFLAGS, "-o", tmp, *SOURCES, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def is_stale() -> bool:
    if not os.path.exists(LIB_PATH):
        return True
    t = os.path.getmtime(LIB_PATH)
    return any(os.path.getmtime(d) > t for d in DEPS)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
