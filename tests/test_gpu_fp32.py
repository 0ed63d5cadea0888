"""The FP32 bulk form (FP32 sweeps and Riemann solves; FP64 pressure recovery,
trouble tests, divergence and RK accumulation) against the FP64 path, which is
itself bit-identical to the reference.  Tolerances are stated per case."""

import numpy as np
import pytest

from fvhoe_amd import initial_conditions as ics
from gpu_helpers import solver_from
from helpers import group, load

pytestmark = pytest.mark.gpu

TRAJ = load("trajectories.npz")
P, O = ("periodic", "periodic"), ("outflow", "outflow")

# L1 bounds on density (mean |rho_fp32 - rho_ref|), measured values are printed
SOD_L1_BOUND = 1e-5
KH_L1_BOUND = 1e-5


def test_sod_fp32_l1():
    g = group(TRAJ, "sod")
    s = solver_from(g["u0"], (256, 1, 1), (3, 0, 0), (O, P, P), riemann_solver="hllc",
                    a_posteriori_slope_limiting=True, bulk_precision="fp32")
    s.ssprk3(0.245)
    w = s.snapshots[-1]["w"]
    l1 = float(np.mean(np.abs(w[0] - g["w_final"][0])))
    print(f"[sod fp32] steps={s.step_count} (fp64: {len(g['timestamps']) - 1}) L1(rho)={l1:.3e}", flush=True)
    assert abs(s.t - 0.245) < 1e-12
    assert l1 < SOD_L1_BOUND


def test_kelvin_helmholtz_fp32_l1():
    n = 256
    u0 = ics.kelvin_helmholtz(n, seed=0)
    kw = dict(riemann_solver="hllc", a_posteriori_slope_limiting=True)
    ref = solver_from(u0, (n, n, 1), (3, 3, 0), (P, P, P), **kw)
    f32 = solver_from(u0, (n, n, 1), (3, 3, 0), (P, P, P), bulk_precision="fp32", **kw)
    ref.ssprk3(n=10)
    f32.ssprk3(n=10)
    d = (f32.u[0] - ref.u[0]).abs().mean().item()
    print(f"[kh256 fp32] 10 steps L1(rho)={d:.3e} t64={ref.t:.6e} t32={f32.t:.6e}", flush=True)
    assert d < KH_L1_BOUND
    assert abs(f32.t - ref.t) / ref.t < 1e-5
