"""MI355X (gfx950) implementation of the acoustic RK4 hot path of the
cubed-sphere solver ankitvb/solar.

The compute path is libsolar_amd.so: hand-written HIP kernels behind a C ABI
(include/solar_amd.h).  This package is the host-side mirror of the
reference's Fortran interface for that path:

    RK4 (timestep.f90:41)           -> Solver.rk4 / solar_rk4
    COMPUTE_RHS (+ BC_RHS, SPONGE)  -> Solver.compute_rhs / solar_compute_rhs
    FILTER_EXP (filter.f90:9)       -> Solver.filter_exp / solar_filter_exp
    SOLAR_DATA / GENERATE_IC        -> background() / generate_ic()

There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.
"""
from .lib import (SolarError, Solver, Ensemble, background, generate_ic, host_tables, load_library,
                  panel_points, RHS_BC, RHS_SPONGE, ALPHA_RK4, BETA_RK4)
from .config import CONFIGS, Config, make_inputs, ensemble_spec, ensemble_member_ic, misfit_problem

__all__ = ["SolarError", "Solver", "Ensemble", "background", "generate_ic", "host_tables", "load_library",
           "panel_points", "RHS_BC", "RHS_SPONGE", "ALPHA_RK4", "BETA_RK4", "CONFIGS", "Config",
           "make_inputs", "ensemble_spec", "ensemble_member_ic", "misfit_problem"]
