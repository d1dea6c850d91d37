"""Build recipe of the gfx950 shared library libmhd_mi355x.so (in-tree).

    python -m magnetohydrodynamics_simulator_amd._build

compiles csrc/*.hip with hipcc for gfx950 only.  Packed-FP32 instruction
selection is disabled for the device code; the host pass warns that it does
not know the feature, which is harmless.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libmhd_mi355x.so")
SOURCES = ["mhd_kernels.hip", "mhd_capi.hip", "mhd_batch.hip", "mhd_adjoint.hip", "mhd_batch_adjoint.hip"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def flags() -> list[str]:
    return [
        "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    ]


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    """Compile every HIP source into LIB; always recompiles."""
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + flags() + (extra or []) + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    errs = [ln for ln in res.stderr.splitlines() if "error" in ln.lower()]
    if errs:
        sys.stderr.write("\n".join(errs) + "\n")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
