"""LibTorch autograd Function of n IMEX steps (csrc/torch_mhd.cpp, built by
``_build.build_torch`` into ``_torch_build/``).

    from magnetohydrodynamics_simulator_amd import torch_ops
    out = torch_ops.imex_steps(fields, eta, dx, dy, dt, nsteps)   # dict of (ny, nx) tensors
    loss(out).backward()                                         # gradients reach fields and eta

``fields`` maps rho, mx, my, E, bxf, byf to compact float64 GPU tensors of
shape (ny, nx) (row-major, bxf on the left face and byf on the bottom face
of each cell, as the library stores them); ``eta`` is a float64 tensor with
one value.  The forward pass is mhd_adjoint_forward (constant dt, implicit
solutions kept, a checkpoint every ``checkpoint_every`` steps); the backward
pass is the discrete adjoint mhd_adjoint_backward.  There is no CPU path:
the extension and the HIP library must both be present.
"""
from __future__ import annotations

import ctypes
import importlib.machinery
import importlib.util
import os

import torch  # noqa: F401  (binds the HIP runtime the library then shares)

from . import _lib

FIELDS = ("rho", "mx", "my", "E", "bxf", "byf")
EXT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_torch_build", "_torch_mhd.so")
_EXT = None


def load():
    """Import the built extension (raises if it is missing)."""
    global _EXT
    if _EXT is None:
        if not os.path.exists(EXT_PATH):
            raise ImportError(f"{EXT_PATH} is missing: build it with python -m "
                              "magnetohydrodynamics_simulator_amd._build --torch")
        if not os.path.exists(_lib.LIB_PATH):
            raise ImportError(f"{_lib.LIB_PATH} is missing")
        ctypes.CDLL(_lib.LIB_PATH, mode=ctypes.RTLD_GLOBAL)   # resolves the extension's dependency
        loader = importlib.machinery.ExtensionFileLoader("_torch_mhd", EXT_PATH)
        spec = importlib.util.spec_from_file_location("_torch_mhd", EXT_PATH, loader=loader)
        mod = importlib.util.module_from_spec(spec)
        loader.exec_module(mod)
        _EXT = mod
    return _EXT


def imex_steps(fields: dict, eta, dx: float, dy: float, dt: float, nsteps: int, checkpoint_every: int = 10,
               cg_tol: float = 2.220446049250313e-16, cg_max_iter: int = 0) -> dict:
    """nsteps IMEX steps of constant dt as a differentiable torch operation;
    returns the final fields as a dict like ``fields``."""
    ext = load()
    if not torch.is_tensor(eta):
        eta = torch.tensor(float(eta), dtype=torch.float64)
    out = ext.imex_steps(*[fields[k] for k in FIELDS], eta, float(dx), float(dy), float(dt), int(nsteps),
                         int(checkpoint_every), float(cg_tol), int(cg_max_iter))
    return dict(zip(FIELDS, out))
