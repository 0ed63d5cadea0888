"""MI355X-native per-RK-stage update for FVHOE's finite-volume Euler solver.

The Python modules keep the reference's module and function names
(fv, hydro, riemann_solvers, slope_limiting, smooth_extrema_detection,
boundary_conditions, solver, ode); every array operation on the stage path
runs in hand-written HIP kernels for gfx950 reached through the C-ABI
library ``libfvhoe_hip.so`` (declared in include/fvhoe_hip.h).
"""

__all__ = ["initial_conditions"]
