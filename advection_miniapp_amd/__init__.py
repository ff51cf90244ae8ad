"""MI355X-native matrix-free DG advection operators and time integrators.

The compute path is libdgops.so (hand-written HIP kernels for gfx950)
behind the C ABI of include/dgops.h; this package is a thin Python mirror
of the reference's operator interfaces on top of it.
"""
from . import _lib
from ._lib import DgError
from .operators import (AdvectionOperation,
                        DiagonallyImplicitRungeKuttaIntegrator,
                        ImplicitRungeKuttaIntegrator,
                        LowStorageRungeKuttaIntegrator, explicit_time_step,
                        implicit_time_step, irk_coefficients, make_config,
                        solver_control)

__all__ = ["AdvectionOperation", "LowStorageRungeKuttaIntegrator",
           "ImplicitRungeKuttaIntegrator",
           "DiagonallyImplicitRungeKuttaIntegrator", "DgError",
           "make_config", "irk_coefficients", "solver_control",
           "explicit_time_step", "implicit_time_step", "_lib"]
