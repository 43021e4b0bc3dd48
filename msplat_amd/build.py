"""Build the gfx950 HIP library ``msplat_amd/_lib/libmsplat_amd.so`` in-tree.

One hipcc invocation compiles the single translation unit
``csrc/msplat_amd.hip`` (which includes every kernel header).  Flags:

* ``-ffp-contract=off`` and correctly rounded fp32 divide/sqrt: the
  per-Gaussian geometry and the sort keys then round exactly like the same
  IEEE operations on the host, which is what makes the sort stage bit-exact.
* ``-munsafe-fp-atomics``: hardware fp32 atomics (torch's device memory is
  coarse grained).
* ``-packed-fp32-ops`` disabled, so hipcc never emits packed fp32 math.

Usage: ``python -m msplat_amd.build`` (or ``__graft_entry__.build()``).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "msplat_amd.hip")
OUT_DIR = os.path.join(HERE, "_lib")
OUT = os.path.join(OUT_DIR, "libmsplat_amd.so")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    """Compile the library; always recompiles.  Returns the .so path."""
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *FLAGS, *(extra or []), SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
