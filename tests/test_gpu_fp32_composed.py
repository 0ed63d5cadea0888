"""The FP32 bulk form of the composed stage (float32 sweeps, face states and
Riemann solves; float64 pressure recovery, CFL maximum, divergence and RK
accumulation) against the oracle's FP32 form bit for bit: stages in 1-D, 2-D
and 3-D (p = 3, 5, 7; LLF and HLLC; periodic, outflow, reflective and free
sides), a 3-D grid large enough for the sliding-window and z-paired sweep
kernels, and SSP-RK3 steps.  The oracle's FP32 form is pinned in
tests/test_oracle_fp32.py."""

import numpy as np
import pytest

from fvhoe_amd import _lib
from fvhoe_amd import fv
from gpu_helpers import assert_bits_equal, solver_from
from oracle import fvhoe_oracle as O

pytestmark = pytest.mark.gpu

P, OUT = ("periodic", "periodic"), ("outflow", "outflow")


def _state(n, seed=0, gamma=1.4):
    rng = np.random.default_rng(seed)
    X, Y, Z = O.fv_uniform_meshgen(n)
    w = np.stack([1 + 0.3 * np.sin(2 * np.pi * (X + 2 * Y)) * np.cos(2 * np.pi * Z) + 0.02 * rng.random(n),
                  0.4 + 0.1 * np.cos(2 * np.pi * Y), -0.3 + 0.2 * X, 0.2 * np.sin(2 * np.pi * Z),
                  1 + 0.2 * X * Y + 0.1 * Z + 0.02 * rng.random(n)])
    return O.compute_conservatives(w, gamma)


CASES = [
    ("3d-p3-hllc-mixed", (12, 10, 8), (3, 3, 3), "hllc", (("outflow", "reflective"), P, ("free", "outflow"))),
    ("3d-p7-llf-periodic", (16, 16, 16), (7, 7, 7), "llf", (P, P, P)),
    ("3d-p2-llf-long-z", (18, 17, 72), (2, 2, 2), "llf", (P, OUT, P)),
    ("2d-p3-hllc", (32, 24, 1), (3, 3, 0), "hllc", (P, ("reflective", "outflow"), P)),
    ("1d-p5-llf", (64, 1, 1), (5, 0, 0), "llf", (OUT, P, P)),
    ("3d-p7-llf-48", (48, 48, 48), (7, 7, 7), "llf", (P, P, P)),
]


def _pair(case):
    name, n, p, rs, bc = case
    u = _state(n)
    s = solver_from(u, n, p, bc, riemann_solver=rs, fused=False, bulk_precision="fp32",
                    a_posteriori_slope_limiting=False)
    sch = O.Scheme(n=n, p=p, gamma=1.4, riemann_solver=rs, bc=bc, bulk="fp32")
    return s, sch, u


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp32_composed_stage_matches_oracle(case):
    s, sch, u = _pair(case)
    assert not s.fused
    dt, dudt = s.hydrodynamics(s.u)
    dto, dudto = O.hydrodynamics(sch, u)
    print(f"[{case[0]}] dt={dt!r} oracle {dto!r}", flush=True)
    assert dt == dto
    assert_bits_equal(dudt.cpu().numpy(), dudto, case[0])


def test_fp32_composed_steps_match_oracle():
    s, sch, u = _pair(CASES[1])
    s.integrator = "ssprk3"
    t = 0.0
    for k in range(3):
        s.take_step()
        dt, u = O.step(sch, u, t, "ssprk3")
        t += dt
        print(f"[fp32 step {k}] t={s.t!r}", flush=True)
        assert s.t == t
        assert_bits_equal(s.u.cpu().numpy(), u, f"step {k}")


@pytest.mark.parametrize("pos", ["c", "t"])
def test_fp32_sweep_kernels_match_oracle(pos):
    """The float32 sliding-window sweeps (64K elements and more) and z-paired sweeps (z >= 64)."""
    rng = np.random.default_rng(5)
    u = rng.standard_normal((5, 21, 19, 75)).astype(np.float32)
    x = _lib.torch().from_numpy(u).to(_lib.device())
    for p in (1, 3, 7):
        mode = "reconstruction" if pos == "t" else "interpolation"
        for axis in (1, 2, 3):
            got = fv.stencil_sweep(x, mode, p, axis, None if pos == "t" else pos)
            assert got.dtype == _lib.torch().float32
            want = (O.transverse_reconstruction(u, p, axis) if pos == "t"
                    else O.conservative_interpolation(u, p, axis, pos))
            assert np.array_equal(got.cpu().numpy().view(np.int32), want.view(np.int32)), (pos, p, axis)
        pair = fv._zpair(x, _lib.STENCIL[pos], p, 1, False)
        assert pair is not None and pair.dtype == _lib.torch().float32
        one = O.transverse_reconstruction if pos == "t" else (lambda a, q, ax: O.conservative_interpolation(a, q, ax, "c"))
        want = one(one(u, p, 1), p, 3)
        assert np.array_equal(pair.cpu().numpy().view(np.int32), want.view(np.int32)), (pos, p, "x-z")


def test_fp32_composed_close_to_fp64():
    n = (24, 24, 24)
    u = _state(n, seed=3)
    kw = dict(riemann_solver="llf", fused=False, a_posteriori_slope_limiting=False)
    s64 = solver_from(u, n, (7, 7, 7), (P, P, P), **kw)
    s32 = solver_from(u, n, (7, 7, 7), (P, P, P), bulk_precision="fp32", **kw)
    for s in (s64, s32):
        s.integrator = "ssprk3"
        for _ in range(4):
            s.take_step()
    l1 = float((s32.u[0] - s64.u[0]).abs().mean().item())
    print(f"[fp32 vs fp64 24^3 p=7, 4 steps] L1(rho)={l1:.3e}", flush=True)
    assert 0 < l1 < 1e-6


def test_fp32_composed_refusals():
    u = _state((12, 10, 8))
    with pytest.raises(NotImplementedError, match="FP32"):
        solver_from(u, (12, 10, 8), (3, 3, 3), (P, P, P), bulk_precision="fp32", a_posteriori_slope_limiting=True)
