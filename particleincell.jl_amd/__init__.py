"""MI355X-native particle-mesh step of ParticleInCell.jl.

The hot path (charge deposition, field gather, the pushers of the CIC, M6,
NGP, TSC and M4 shapes, and the periodic FFT Poisson solve) runs in
libpicmi355x.so: hand-written HIP kernels for gfx950 plus rocFFT behind the
C ABI of include/pic_mi355x.h.
This package is the host side: ctypes bindings (_lib), the reference's API
(api), the Landau quiet start (sampling) and the damping-rate fit
(diagnostics).
"""
from . import _lib
from .api import (CloudInCell, Context, F90, M4, M6, MeshFields2D, NearestGridPoint,
                  ParticleGroup, Particles, TriangularShapeCloud,
                  Poisson2D, Simulation, TwoDGrid, TwoDPoissonPeriodic, compute_rho,
                  compute_rho_m6_, energy, interpol_eb_m6_, interpolate_cic_, push_v_, push_x_,
                  solve_)
from . import sampling, diagnostics

__all__ = [
    "CloudInCell", "Context", "F90", "M4", "M6", "MeshFields2D", "NearestGridPoint",
    "ParticleGroup", "Particles", "TriangularShapeCloud",
    "Poisson2D", "Simulation", "TwoDGrid", "TwoDPoissonPeriodic", "compute_rho",
    "compute_rho_m6_", "energy", "interpol_eb_m6_", "interpolate_cic_", "push_v_", "push_x_",
    "solve_", "sampling", "diagnostics",
]
