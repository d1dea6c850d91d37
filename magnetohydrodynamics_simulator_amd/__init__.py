"""MI355X-native (gfx950, HIP) FP64 step of 2D resistive MHD.

Drop-in for the hot path of RikibabManos/Magnetohydrodynamics_Simulator:
MUSCL-minmod reconstruction, HLL x/y fluxes, upwind constrained transport and
the ARS(2,2,2) IMEX step with a matrix-free Jacobi-PCG implicit resistive
solve.  The compute runs in libmhd_mi355x.so (C ABI: include/mhd_mi355x.h);
this package is its Python host binding.
"""
from .solver import GAMMA_ARS, Batch, Context, Grid, operator_alpha  # noqa: F401
from . import setups  # noqa: F401

__all__ = ["Batch", "Context", "Grid", "GAMMA_ARS", "operator_alpha", "setups"]
