"""Parameters of the golden stage cases (shared by make_golden.py and the tests).

Each case is (name, n, p, bc, warm_steps, seed, u0, kwargs): EulerSolver
keyword arguments of the reference (fvhoe/solver.py:42-84) for one stage
recorded after `warm_steps` SSP-RK3 steps from a seeded random state.
"""

from fvhoe_amd import initial_conditions as ics

STAGE_CASES = []


def stage_case(name, n, p, bc, warm_steps=0, seed=0, u0=None, **kw):
    STAGE_CASES.append((name, n, p, bc, warm_steps, seed, u0, kw))


def build_stage_cases():
    P = ("periodic", "periodic")
    O = ("outflow", "outflow")
    R = ("reflective", "reflective")
    F = ("free", "free")
    apost = dict(a_posteriori_slope_limiting=True)
    # 1-D, every p, both solvers, every BC kind
    for p in range(0, 9):
        stage_case(f"1d_p{p}_hllc_apost", (48, 1, 1), (p, 0, 0), (O, P, P), warm_steps=6, seed=p,
                   riemann_solver="hllc", **apost)
    for p in (1, 3, 7):
        stage_case(f"1d_p{p}_llf", (40, 1, 1), (p, 0, 0), (P, P, P), warm_steps=3, seed=10 + p,
                   riemann_solver="llf")
    stage_case("1d_p2_upwind", (40, 1, 1), (2, 0, 0), (P, P, P), seed=21, riemann_solver="advection_upwind")
    stage_case("1d_p3_reflective", (48, 1, 1), (3, 0, 0), (R, P, P), warm_steps=5, seed=22, **apost)
    stage_case("1d_p4_free", (48, 1, 1), (4, 0, 0), (F, P, P), warm_steps=5, seed=23, **apost)
    stage_case("1d_p3_y", (1, 48, 1), (0, 3, 0), (P, O, P), warm_steps=5, seed=24, **apost)
    stage_case("1d_p3_z", (1, 1, 48), (0, 0, 3), (P, P, O), warm_steps=5, seed=25, **apost)
    stage_case("1d_p5_mixed_bc", (48, 1, 1), (5, 0, 0), (("outflow", "reflective"), P, P),
               warm_steps=4, seed=26, **apost)
    # 2-D
    stage_case("2d_p3_hllc_apost", (24, 20, 1), (3, 3, 0), (P, P, P), warm_steps=4, seed=30, **apost)
    stage_case("2d_p3_outflow_refl", (24, 20, 1), (3, 3, 0), (O, R, P), warm_steps=4, seed=31, **apost)
    stage_case("2d_p2_free", (20, 18, 1), (2, 2, 0), (F, O, P), warm_steps=3, seed=32, **apost)
    stage_case("2d_p5_hllc_apost", (26, 24, 1), (5, 5, 0), (P, P, P), warm_steps=3, seed=33, **apost)
    stage_case("2d_p8_llf", (24, 20, 1), (8, 8, 0), (P, P, P), warm_steps=2, seed=34, riemann_solver="llf")
    stage_case("2d_p1_apost", (20, 18, 1), (1, 1, 0), (O, O, P), warm_steps=3, seed=35, **apost)
    stage_case("2d_p3_p2_mixed", (22, 20, 1), (3, 2, 0), (P, O, P), warm_steps=3, seed=36, **apost)
    stage_case("2d_p3_xz", (20, 1, 18), (3, 0, 3), (P, P, O), warm_steps=3, seed=37, **apost)
    stage_case("2d_p3_yz", (1, 20, 18), (0, 3, 3), (P, O, P), warm_steps=3, seed=38, **apost)
    stage_case("2d_force_minmod", (20, 18, 1), (3, 3, 0), (P, O, P), warm_steps=2, seed=39,
               force_trouble=True, **apost)
    stage_case("2d_force_moncen", (20, 18, 1), (3, 3, 0), (P, O, P), warm_steps=2, seed=40,
               force_trouble=True, slope_limiter="moncen", **apost)
    stage_case("2d_convex", (22, 20, 1), (3, 3, 0), (P, O, P), warm_steps=3, seed=41, convex=True, **apost)
    stage_case("2d_nad_local_abs", (22, 20, 1), (3, 3, 0), (P, P, P), warm_steps=3, seed=42,
               NAD=1e-3, NAD_mode="local", NAD_range="absolute", **apost)
    stage_case("2d_nad_local_rel_vars", (22, 20, 1), (3, 3, 0), (P, P, P), warm_steps=3, seed=43,
               NAD=1e-4, NAD_mode="local", NAD_vars=("rho", "P"), SED=False, **apost)
    stage_case("2d_nad_global_abs_pad", (22, 20, 1), (3, 3, 0), (P, P, P), warm_steps=3, seed=44,
               NAD=1e-5, NAD_range="absolute", PAD={"P": (0.0, 3.0), "rho": (0.9, 10.0)}, **apost)
    # 3-D
    stage_case("3d_p3_hllc_apost", (10, 9, 8), (3, 3, 3), (P, O, P), warm_steps=2, seed=50, **apost)
    stage_case("3d_p2_refl", (10, 9, 8), (2, 2, 2), (R, P, F), warm_steps=2, seed=51, **apost)
    stage_case("3d_p7_llf", (16, 16, 16), (7, 7, 7), (P, P, P), seed=52, riemann_solver="llf",
               u0=ics.density_wave_3d(16, 7))


# initial conditions: (key, class name, ctor kwargs, n, p, domain); small meshes, every IC class
IC_CASES = [
    ("sinus_3d", "Sinus", dict(dims="xyz", rho_min_max=(0.5, 1.5), vx=1, vy=1, vz=1), (8, 6, 5), (7, 2, 1),
     ((0, 1), (0, 1), (0, 1))),
    ("square_xy", "Square", dict(dims="xy"), (16, 12, 1), (2, 2, 0), ((0, 1), (0, 1), (0, 1))),
    ("slotted_disk", "SlottedDisk", {}, (20, 20, 1), (3, 3, 0), ((0, 1), (0, 1), (0, 1))),
    ("shock1d_y", "Shock1D", dict(dim="y"), (1, 32, 1), (0, 3, 0), ((0, 1), (0, 1), (0, 1))),
    ("double_shock", "DoubleShock1D", {}, (40, 1, 1), (4, 0, 0), ((0, 1), (0, 1), (0, 1))),
    ("shu_osher_z", "ShuOsher1D", dict(dim="z"), (1, 1, 32), (0, 0, 3), ((0, 1), (0, 1), (0, 10))),
    ("kelvin_helmholtz", "KelvinHelmholtz2D", {}, (24, 24, 1), (3, 3, 0), ((0, 1), (0, 1), (0, 1))),
    ("double_mach", "DoubleMachReflection2D", dict(gamma=1.4), (32, 8, 1), (1, 1, 0), ((0, 4), (0, 1), (0, 1))),
    ("shock_tube_ball", "ShockTube", dict(dims="xy", center=(0.5, 0.5), radius=0.3), (16, 16, 1), (2, 2, 0),
     ((0, 1), (0, 1), (0, 1))),
    ("shock_tube_cube", "ShockTube", dict(mode="cube", x_cube=(0.2, 0.6), y_cube=(0.3, 0.7)), (16, 16, 1),
     (2, 2, 0), ((0, 1), (0, 1), (0, 1))),
    ("sedov_center", "SedovBlast", dict(mode="center", dims="xy"), (16, 16, 1), (0, 0, 0),
     ((0, 1), (0, 1), (0, 1))),
    ("sedov_corner", "SedovBlast", dict(mode="corner", dims="xy"), (12, 12, 1), (2, 2, 0),
     ((0, 1), (0, 1), (0, 1))),
    ("athena_blast", "AthenaBlast", {}, (16, 24, 1), (2, 2, 0), ((-0.5, 0.5), (-0.75, 0.75), (0, 1))),
]
