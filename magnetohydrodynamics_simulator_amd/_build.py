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
SOURCES = ["mhd_kernels.hip", "mhd_capi.hip", "mhd_batch.hip", "mhd_adjoint.hip", "mhd_batch_adjoint.hip", "mhd_io.hip",
           "mhd_resident.hip", "mhd_setups.hip", "mhd_noalloc.hip", "mhd_snapshots.hip"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def flags() -> list[str]:
    """Compile flags of every source (gfx950 device code, host PIC)."""
    return [
        "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics",
        "-Xarch_host", "-fvisibility=hidden",     # export only the C ABI (MHD_API)
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    ]


def _run(cmd: list[str], verbose: bool) -> subprocess.CompletedProcess:
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    """Compile every HIP source into LIB; always recompiles.  The sources are
    compiled to objects in parallel (each carries its own gfx950 code object)
    and linked into one shared library."""
    from concurrent.futures import ThreadPoolExecutor

    obj_dir = os.path.join(PKG, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[hipcc()] + flags() + (extra or []) + ["-c", os.path.join(CSRC, s), "-o", o]
            for s, o in zip(SOURCES, objs)]
    jobs = max(1, min(len(cmds), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda c: _run(c, verbose), cmds))
    for res in results:
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
        errs = [ln for ln in res.stderr.splitlines() if "error" in ln.lower()]
        if errs:
            sys.stderr.write("\n".join(errs) + "\n")
    tmp = LIB + ".tmp"
    res = _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,libmhd_mi355x.so"] + objs +
               ["-o", tmp], verbose)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc link failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


TORCH_SRC = os.path.join(CSRC, "torch_mhd.cpp")
TORCH_BUILD = os.path.join(PKG, "_torch_build")
TORCH_NAME = "_torch_mhd"


def build_torch(verbose: bool = False) -> str:
    """Compile the LibTorch autograd Function (csrc/torch_mhd.cpp) against the
    torch wheel's headers into _torch_build/_torch_mhd*.so, linked to LIB.
    Host C++ only (the GPU work is the library's); always recompiles."""
    import ctypes

    from torch.utils import cpp_extension

    # the extension's dependency on libmhd_mi355x.so resolves to this copy
    ctypes.CDLL(LIB, mode=ctypes.RTLD_GLOBAL)
    os.makedirs(TORCH_BUILD, exist_ok=True)
    for f in os.listdir(TORCH_BUILD):          # force a fresh build
        if f.endswith((".o", ".so", ".ninja_deps", ".ninja_log")) or f == "build.ninja":
            os.remove(os.path.join(TORCH_BUILD, f))
    cpp_extension.load(
        name=TORCH_NAME, sources=[TORCH_SRC], build_directory=TORCH_BUILD, verbose=verbose,
        extra_include_paths=[os.path.join(ROOT, "include")],
        extra_cflags=["-O2"], extra_ldflags=[f"-L{PKG}", "-lmhd_mi355x"], with_cuda=False)
    so = [f for f in os.listdir(TORCH_BUILD) if f.startswith(TORCH_NAME) and f.endswith(".so")]
    if not so:
        raise RuntimeError("the LibTorch extension was not produced")
    return os.path.join(TORCH_BUILD, so[0])


if __name__ == "__main__":
    print(build(verbose=True))
    if "--torch" in sys.argv:
        print(build_torch(verbose=True))
