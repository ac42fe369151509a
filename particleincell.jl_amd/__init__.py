"""MI355X-native particle-mesh step of ParticleInCell.jl.

The hot path (charge deposition, field gather, the pushers of the CIC, M6,
NGP, TSC and M4 shapes, the periodic FFT Poisson solve, and the FDTD
Vlasov-Maxwell loop with its CIC current deposit, the uniformly accurate
(UA) step) runs in
libpicmi355x.so: hand-written HIP kernels for gfx950 plus rocFFT behind the
C ABI of include/pic_mi355x.h.
This package is the host side: ctypes bindings (_lib), the reference's API
(api), the Landau quiet start (sampling) and the damping-rate fit
(diagnostics).
"""
from . import _lib
from .api import (FDTD, CloudInCell, Context, F90, M4, M6, MeshFields2D, NearestGridPoint,
                  ParticleGroup, Particles, TriangularShapeCloud,
                  Poisson2D, Simulation, TwoDGrid, TwoDPoissonPeriodic, ampere_maxwell_,
                  compute_current_, compute_energy, compute_rho, compute_rho_m6_, energy,
                  faraday_, interpol_eb_m6_, interpolate_cic_, push_v_, push_x_, solve_,
                  update_fields_, UA, UASimulation, preparation_, update_particles_e_,
                  compute_f_, ua_step_, update_particles_x_, compute_v_, read_particles,
                  write_particles, operator_t_)
from . import sampling, diagnostics

__all__ = [
    "FDTD", "ampere_maxwell_", "compute_current_", "compute_energy", "faraday_",
    "update_fields_", "UA", "UASimulation", "preparation_", "update_particles_e_", "compute_f_",
    "ua_step_", "update_particles_x_", "compute_v_", "read_particles", "write_particles",
    "operator_t_",
    "CloudInCell", "Context", "F90", "M4", "M6", "MeshFields2D", "NearestGridPoint",
    "ParticleGroup", "Particles", "TriangularShapeCloud",
    "Poisson2D", "Simulation", "TwoDGrid", "TwoDPoissonPeriodic", "compute_rho",
    "compute_rho_m6_", "energy", "interpol_eb_m6_", "interpolate_cic_", "push_v_", "push_x_",
    "solve_", "sampling", "diagnostics",
]
