"""Dirichlet and "ic" boundaries on the GPU (fvhoe/boundary_conditions.py:18-57,
247-348): the device fill against the oracle's (itself pinned in
tests/test_oracle_boundaries.py), the fused and the composed stage against
the oracle's stage bit for bit, time-dependent values refreshed per step, and
the reference's "ic" == constant Dirichlet case (tests/test_boundary_conditions.py:147-230)."""

import numpy as np
import pytest

from fvhoe_amd import _lib
from fvhoe_amd.boundary_conditions import BoundaryCondition
from fvhoe_amd.fv import fv_uniform_meshgen
from fvhoe_amd.initial_conditions import InitialCondition, ShuOsher1D
from fvhoe_amd.solver import EulerSolver
from gpu_helpers import assert_bits_equal
from oracle import fvhoe_oracle as O

pytestmark = pytest.mark.gpu


def _f(x, y, z, t):
    d = np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z) * np.sin(2 * np.pi * t)
    return np.stack([d, 3 * d, 4 * d, 5 * d, 2 * d])


def _g(x, y, z, t=None):  # a positive state for the stage tests (no t: usable as an "ic"-like value)
    return np.stack([1.2 + 0.3 * np.sin(3 * x + y), 0.4 * np.cos(2 * y), -0.3 + 0.1 * x, 0.0 * x,
                     3.0 + 0.5 * np.sin(x - 2 * y)])


def _gt(x, y, z, t):
    out = _g(x, y, z)
    out[0] += 0.2 * t
    out[4] += t
    return out


class Smooth2D(InitialCondition):
    """rho, v, P smooth and positive; pointwise primitives."""

    def base_ic(self):
        def f(x, y, z):
            out = np.empty((5,) + x.shape)
            out[0] = 1.0 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
            out[1] = 0.6 + 0.1 * np.sin(2 * np.pi * y)
            out[2] = -0.4 + 0.2 * x
            out[3] = 0.0
            out[4] = 1.0 + 0.2 * x * y + 0.5 * (x > 0.55) * (y > 0.4)
            return out

        return f


def test_device_fill_f_of_xyzt(N=32, gw=8, t=0.5):
    h = 1 / N
    Xa, Ya, Za = fv_uniform_meshgen((N + 2 * gw,) * 3, *((-h * gw, 1 + h * gw),) * 3)
    (X, Y, Z), slabs = fv_uniform_meshgen((N, N, N), slab_thickness=(gw, gw, gw))
    bc = BoundaryCondition(x="dirichlet", y="dirichlet", z="dirichlet", x_value=_f, y_value=_f, z_value=_f,
                           slab_coords=slabs)
    out = bc.apply(_f(X, Y, Z, t), gw=(gw, gw, gw), t=t)
    assert np.array_equal(out, _f(Xa, Ya, Za, t))


@pytest.mark.parametrize("kinds", [
    (("dirichlet", "outflow"), ("reflective", "ic"), ("periodic", "periodic")),
    (("free", "dirichlet"), ("dirichlet", "dirichlet"), ("outflow", "reflective")),
    (("reflective", "ic"), ("periodic", "periodic"), ("dirichlet", "free")),
], ids=["x-dir", "y-both", "z-dir"])
def test_device_fill_matches_oracle(kinds):
    n, gw = (9, 7, 5), (3, 2, 2)
    rng = np.random.default_rng(1)
    u = rng.random((5, *n))
    (_, _, _), slabs = fv_uniform_meshgen(n, slab_thickness=(4, 4, 4))
    c = np.array([0.5, -1.0, 2.0, 0.25, 7.0])
    pick = {"dirichlet": _gt, "ic": _g}
    vals = tuple(tuple(c if (k == "dirichlet" and (a + j) % 2) else pick.get(k) for j, k in enumerate(kk))
                 for a, kk in enumerate(kinds))
    bc = BoundaryCondition(x=kinds[0], y=kinds[1], z=kinds[2], x_value=vals[0], y_value=vals[1], z_value=vals[2],
                           slab_coords=slabs)
    got = bc.apply(u, gw, t=0.7)
    want = O.apply_bc(u, gw, kinds, vals, slabs, 0.7)
    print(f"[fill {kinds}]", flush=True)
    assert_bits_equal(got, want, "fill")


def _solver(n, p, bc, fused=None, **kw):
    kw.setdefault("gamma", 1.4)
    return EulerSolver(w0=Smooth2D(), nx=n[0], ny=n[1], nz=n[2], px=p[0], py=p[1], pz=p[2], bc=bc,
                       riemann_solver="hllc", a_posteriori_slope_limiting=True, progress_bar=False, fused=fused, **kw)


def _scheme(s, kinds, vals):
    return O.Scheme(n=s.n, p=s.p, gamma=s.gamma, a_posteriori=True, bc=kinds, bc_values=vals,
                    riemann_solver="hllc", slope_limiter="minmod")


STAGE_CASES = [
    ("2d-p3-x", (48, 40, 1), (3, 3, 0), (("dirichlet", "ic"), ("outflow", "outflow"), (None, None))),
    ("2d-p3-xy", (40, 36, 1), (3, 3, 0), (("ic", "reflective"), ("dirichlet", "dirichlet"), (None, None))),
    ("2d-p2-y", (32, 40, 1), (2, 2, 0), (("periodic", "periodic"), ("ic", "dirichlet"), (None, None))),
    ("1d-p5", (64, 1, 1), (5, 0, 0), (("ic", "dirichlet"), (None, None), (None, None))),
    ("3d-p2", (12, 10, 8), (2, 2, 2), (("dirichlet", "outflow"), ("periodic", "periodic"), ("ic", "free"))),
]


@pytest.mark.parametrize("case", STAGE_CASES, ids=[c[0] for c in STAGE_CASES])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "composed"])
def test_stage_matches_oracle(case, fused):
    name, n, p, kinds = case
    c = np.array([1.1, 0.2, -0.1, 0.0, 2.5])
    s0 = _solver(n, p, BoundaryCondition())  # for the pointwise IC
    vals = tuple(tuple({"dirichlet": (_gt if j == 0 else c), "ic": s0.u0}.get(k) for j, k in enumerate(kk))
                 for kk in kinds)
    bc = BoundaryCondition(x=kinds[0], y=kinds[1], z=kinds[2] if kinds[2] != (None, None) else "periodic",
                           x_value=vals[0], y_value=vals[1], z_value=vals[2])
    s = _solver(n, p, bc, fused=None if fused else False)
    if fused and not s.fused:
        pytest.skip("the fused stage does not cover this configuration")
    sch = _scheme(s, tuple(k if k != (None, None) else ("periodic", "periodic") for k in kinds), vals)
    u0 = s.u.cpu().numpy()
    dt, dudt = s.hydrodynamics(s.u)
    dto, dudto = O.hydrodynamics(sch, u0)
    print(f"[{name} fused={s.fused}] dt={dt!r}", flush=True)
    assert dt == dto
    assert_bits_equal(dudt.cpu().numpy(), dudto, name)


def test_time_dependent_dirichlet_steps_match_oracle():
    """Values that depend on t are re-evaluated at each step's start time (fvhoe/solver.py:784)."""
    n, p = (40, 32, 1), (3, 3, 0)
    c = np.array([1.1, 0.2, -0.1, 0.0, 2.5])
    kinds = (("dirichlet", "outflow"), ("dirichlet", "reflective"), ("periodic", "periodic"))
    vals = ((_gt, None), (c, None), (None, None))
    bc = BoundaryCondition(x=kinds[0], y=kinds[1], x_value=vals[0], y_value=vals[1])
    s = _solver(n, p, bc)
    assert s.fused
    s.integrator = "ssprk3"
    sch = _scheme(s, kinds, vals)
    u = s.u.cpu().numpy()
    t = 0.0
    for k in range(3):
        s.take_step()
        dt, u = O.step(sch, u, t, "ssprk3")
        t += dt
        assert s.t == t
        assert_bits_equal(s.u.cpu().numpy(), u, f"step {k}")
    with pytest.raises(NotImplementedError, match="host clock"):
        s.advance(1)


@pytest.mark.parametrize("dim,p", [("x", 0), ("x", 1), ("x", 2), ("x", 3), ("x", 8), ("y", 3), ("z", 1)])
def test_ic_equals_constant_dirichlet_shu_osher(dim, p, N=100, gamma=1.4, T=1.8):
    """tests/test_boundary_conditions.py:147-230 (solver.run picks the RK order from p)."""
    cfg = dict(w0=ShuOsher1D(dim=dim), x=(0, 10) if dim == "x" else (0, 1), y=(0, 10) if dim == "y" else (0, 1),
               z=(0, 10) if dim == "z" else (0, 1), nx=N if dim == "x" else 1, ny=N if dim == "y" else 1,
               nz=N if dim == "z" else 1, px=p if dim == "x" else 0, py=p if dim == "y" else 0,
               pz=p if dim == "z" else 0, riemann_solver="hllc", gamma=gamma, a_posteriori_slope_limiting=p > 0,
               slope_limiter="minmod", progress_bar=False)
    arrs = (np.array([3.857143, 3.857143 * 2.629369 if dim == "x" else 0, 3.857143 * 2.629369 if dim == "y" else 0,
                      3.857143 * 2.629369 if dim == "z" else 0,
                      10.33333 / (gamma - 1) + 0.5 * 3.857143 * 2.629369 * 2.629369]),
            np.array([1 + 0.2 * np.sin(5 * 5), 0, 0, 0, 1 / (gamma - 1)]))
    sd = EulerSolver(bc=BoundaryCondition(x="dirichlet" if dim == "x" else None, y="dirichlet" if dim == "y" else None,
                                          z="dirichlet" if dim == "z" else None,
                                          x_value=arrs if dim == "x" else None, y_value=arrs if dim == "y" else None,
                                          z_value=arrs if dim == "z" else None), **cfg)
    si = EulerSolver(bc=BoundaryCondition(x="ic" if dim == "x" else None, y="ic" if dim == "y" else None,
                                          z="ic" if dim == "z" else None), **cfg)
    print(f"[ic {dim} p={p}] fused={sd.fused}", flush=True)
    sd.run(T)
    si.run(T)
    Pd = sd.snapshots[-1]["w"][4]
    Pi = si.snapshots[-1]["w"][4]
    assert O.l1err(Pd, Pi) < 1e-14


def test_slab_boundaries_refuse_the_predictor_and_the_ensemble_checks_its_inputs():
    """The predictor's features map ghosts into the interior, so slab sides refuse it; the ensemble takes
    "ic" sides only with every instance's pointwise IC, and constant Dirichlet values only."""
    from fvhoe_amd.ensemble import Ensemble
    from fvhoe_amd.predictor import TroubledCellPredictor

    bc = BoundaryCondition(x="ic", y="outflow")
    with pytest.raises(NotImplementedError):
        _solver((32, 32, 1), (3, 3, 0), bc, predictor=TroubledCellPredictor.load())
    with pytest.raises(ValueError, match="pointwise initial condition"):
        Ensemble([np.ones((5, 32))], bc=BoundaryCondition(x="ic"))
    with pytest.raises(NotImplementedError, match="constant Dirichlet"):
        Ensemble([np.ones((5, 32))], bc=BoundaryCondition(x="dirichlet", x_value=(lambda x, y, z, t: x, None)))
