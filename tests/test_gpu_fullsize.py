"""Parity at the workloads' full sizes, where the large-grid kernels run
(XCD-ordered tiles, sliding windows, 32-bit offset limits): a fused FP64
stage of Kelvin-Helmholtz 2048^2 and of the four-quadrant problem at 1024^2,
a composed FP64 stage of the 3-D density wave at 128^3 p=7 (LLF), each
against the oracle bit for bit (dt, dudt, trouble flags and violation
magnitudes); the 512-instance Shu-Osher ensemble against single fused runs of
sampled instances; and the FP32 bulk form at 2048^2 against FP64 within the
stated L1 bound.  The oracle stages take seconds each on one host core."""

import time

import numpy as np
import pytest

import oracle as O
from fvhoe_amd import initial_conditions as ics
from fvhoe_amd.ensemble import shu_osher_ensemble_solver
from gpu_helpers import assert_bits_equal, solver_from

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
P, OUT = ("periodic", "periodic"), ("outflow", "outflow")

KH2048_FP32_L1_BOUND = 1e-5  # L1(rho) FP32 vs FP64 after 10 SSP-RK3 steps at 2048^2


def _stage_vs_oracle(s, scheme, u, tag):
    """One stage of s on u (NumPy) against the oracle's, bit for bit."""
    t0 = time.time()
    dt, dudt = s.hydrodynamics(u)
    tg = time.time() - t0
    log = O.StageLog()
    t0 = time.time()
    dto, dudto = O.hydrodynamics(scheme, u, log)
    print(f"[{tag}] dt={dt!r} gpu {tg:.2f}s oracle {time.time() - t0:.1f}s", flush=True)
    assert dt == dto, (dt, dto)
    assert_bits_equal(dudt, dudto, tag + " dudt")
    if scheme.a_posteriori:
        flags = s._trouble.cpu().numpy().astype(bool)
        want = np.asarray(log.trouble).astype(bool).reshape(flags.shape)
        print(f"[{tag}] troubled cells {int(want.sum())}", flush=True)
        assert np.array_equal(flags, want), (tag, int((flags != want).sum()))
        # a solver whose log was reset before this stage: the log holds 0 + this stage's magnitudes
        assert_bits_equal(s.log["mean NAD mag"].cpu().numpy(), log.NAD_mag, tag + " NAD magnitude")
        assert_bits_equal(s.log["mean PAD mag"].cpu().numpy(), log.PAD_mag, tag + " PAD magnitude")
    return dt, dudt


def _fresh_log(s):
    s.log_troubles(reset=True)  # the next stage starts a new running sum (the reference assigns)


def _kh2048():
    n = 2048
    u0 = ics.kelvin_helmholtz(n, seed=0)
    s = solver_from(u0, (n, n, 1), (3, 3, 0), (P, P, P), riemann_solver="hllc", a_posteriori_slope_limiting=True)
    assert s.fused
    return u0, s, O.Scheme(n=(n, n, 1), p=(3, 3, 0), gamma=1.4, a_posteriori=True)


def test_kelvin_helmholtz_2048_fused_stage_1():
    u0, s, sch = _kh2048()
    _stage_vs_oracle(s, sch, u0, "kh2048 stage 1")


def test_kelvin_helmholtz_2048_fused_stage_2():
    """The second stage's input of an SSP-RK3 step, formed as the reference forms it (fvhoe/ode.py:276)
    from the first stage (the GPU's, bit-identical to the oracle's by the test above)."""
    u0, s, sch = _kh2048()
    dt, dudt = s.hydrodynamics(u0)
    _fresh_log(s)
    _stage_vs_oracle(s, sch, u0 + dt * dudt, "kh2048 stage 2")


def test_four_quadrants_1024_fused_stage():
    n = 1024
    u0 = ics.four_quadrants(n, seed=200)
    s = solver_from(u0, (n, n, 1), (3, 3, 0), (OUT, OUT, P), riemann_solver="hllc", a_posteriori_slope_limiting=True)
    assert s.fused
    sch = O.Scheme(n=(n, n, 1), p=(3, 3, 0), gamma=1.4, a_posteriori=True, bc=(OUT, OUT, P))
    dt, dudt = _stage_vs_oracle(s, sch, u0, "quad1024 stage 1")
    _fresh_log(s)
    _stage_vs_oracle(s, sch, u0 + dt * dudt, "quad1024 stage 2")


def test_density_wave_128_p7_composed_stage():
    n = 128
    u0 = ics.density_wave_3d(n, 7)
    s = solver_from(u0, (n, n, n), (7, 7, 7), (P, P, P), riemann_solver="llf")
    assert not s.fused
    sch = O.Scheme(n=(n, n, n), p=(7, 7, 7), gamma=1.4, riemann_solver="llf")
    _stage_vs_oracle(s, sch, u0, "wave128 p7")


@pytest.mark.parametrize("boundary", ["ic", "outflow"])
def test_ensemble_512_against_single_runs(boundary):
    """Sampled instances of the 512-run ensemble against single fused runs with the same boundaries: "ic"
    ghosts (the initial condition at the ghost centres, as the reference runs Shu-Osher) or outflow."""
    import torch

    from fvhoe_amd.boundary_conditions import BoundaryCondition
    from fvhoe_amd.solver import EulerSolver

    e = shu_osher_ensemble_solver(512, seed=0, boundary=boundary)
    insts = ics.shu_osher_ensemble(512, seed=0)
    u0 = [e.state(i).copy() for i in range(e.count)]
    steps = 4
    for _ in range(steps):
        e.step(1e9)
    for i in (0, 17, 100, 255, 400, 511):
        n = int(e.n[i])
        if boundary == "ic":
            # the single run's "ic" ghosts come from its pointwise IC; its cells start from the ensemble's
            # cell-centre values
            s = EulerSolver(w0=ics.ShuOsherInstance(eps=insts[i][1], k=insts[i][2]), nx=n, px=5, x=(-5, 5),
                            gamma=1.4, riemann_solver="hllc", a_posteriori_slope_limiting=True, progress_bar=False,
                            bc=BoundaryCondition(x="ic"))
            s.u.copy_(torch.from_numpy(u0[i].reshape(s.u.shape)))
        else:
            s = solver_from(u0[i].reshape(5, n, 1, 1), (n, 1, 1), (5, 0, 0), (OUT, P, P), x=(-5, 5),
                            riemann_solver="hllc", a_posteriori_slope_limiting=True)
        assert s.fused
        s.integrator = "ssprk3"
        for _ in range(steps):
            s.take_step()
        print(f"[ensemble 512] instance {i} n={n} t={s.t!r}", flush=True)
        assert s.t == e.times()[i]
        assert_bits_equal(e.state(i), s.u.cpu().numpy().reshape(5, n), f"instance {i}")


def test_kelvin_helmholtz_2048_fp32_l1():
    n = 2048
    u0 = ics.kelvin_helmholtz(n, seed=0)
    kw = dict(riemann_solver="hllc", a_posteriori_slope_limiting=True)
    ref = solver_from(u0, (n, n, 1), (3, 3, 0), (P, P, P), **kw)
    f32 = solver_from(u0, (n, n, 1), (3, 3, 0), (P, P, P), bulk_precision="fp32", **kw)
    ref.integrator = f32.integrator = "ssprk3"
    ref.advance(10)
    f32.advance(10)
    d = (f32.u[0] - ref.u[0]).abs().mean().item()
    print(f"[kh2048 fp32] 10 steps L1(rho)={d:.3e}", flush=True)
    assert d < KH2048_FP32_L1_BOUND
