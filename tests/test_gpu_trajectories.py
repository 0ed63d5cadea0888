"""Multi-step parity on the GPU against trajectories the reference wrote:
Sod N=256 to t=0.245, Kelvin-Helmholtz 256^2 (trouble bitmaps of every
stage), 3-D density wave 16^3 p=7, four quadrants 64^2, Shu-Osher N=512 p=5
and the Euler / SSP-RK2 / RK4 steppers."""

import hashlib

import numpy as np
import pytest

from fvhoe_amd import initial_conditions as ics
from gpu_helpers import assert_bits_equal, solver_from
from helpers import group, load

pytestmark = pytest.mark.gpu

TRAJ = load("trajectories.npz")
P, O = ("periodic", "periodic"), ("outflow", "outflow")


def _record_flags(s, sink):
    """Copy every stage's flags to the host (so the steps run one by one, not as graph replays)."""
    orig = s._log_troubles_device

    def rec(flags, gate=None):
        sink.append(flags.cpu().numpy().astype(bool))
        orig(flags, gate=gate)

    s._log_troubles_device = rec
    s._chunkable = lambda: False


@pytest.mark.parametrize("fused", [True, False])
def test_sod_to_final_time(fused):
    g = group(TRAJ, "sod")
    s = solver_from(g["u0"], (256, 1, 1), (3, 0, 0), (O, P, P), fused=fused, riemann_solver="hllc",
                    a_posteriori_slope_limiting=True)
    assert s.fused == fused
    s.ssprk3(0.245)
    print(f"[sod fused={fused}] steps={s.step_count} t={s.t!r}", flush=True)
    assert s.t == float(g["t"])
    assert s.step_count == len(g["timestamps"]) - 1
    assert np.array_equal(np.array(s.timestamps), g["timestamps"])
    assert_bits_equal(s.u.cpu().numpy(), g["u_final"], "u")
    assert_bits_equal(s.snapshots[-1]["w"], g["w_final"], "w snapshot")


def test_kelvin_helmholtz_256_flags_and_state():
    g = group(TRAJ, "kh256")
    u0 = ics.kelvin_helmholtz(256, seed=0)
    assert hashlib.sha256(u0.tobytes()).hexdigest() == str(g["u0_sha"])
    s = solver_from(u0, (256, 256, 1), (3, 3, 0), (P, P, P), riemann_solver="hllc",
                    a_posteriori_slope_limiting=True)
    assert s.fused
    flags = []
    _record_flags(s, flags)
    s.ssprk3(n=3)
    bits = np.stack([np.packbits(f.ravel()) for f in flags])
    print(f"[kh256] troubled per stage {[int(f.sum()) for f in flags]}", flush=True)
    assert np.array_equal(bits, g["trouble_bits"])
    uf = s.u.cpu().numpy()
    assert_bits_equal(uf[:, ::37, :, 0], g["u_final_rows"], "rows")
    assert hashlib.sha256(np.ascontiguousarray(uf).tobytes()).hexdigest() == str(g["u_final_sha"])
    assert np.array_equal(np.array(s.timestamps), g["timestamps"])


def test_density_wave_3d_p7():
    g = group(TRAJ, "wave3d")
    s = solver_from(g["u0"], (16, 16, 16), (7, 7, 7), (P, P, P), riemann_solver="llf")
    s.ssprk3(n=2)
    assert_bits_equal(s.u.cpu().numpy(), g["u_final"], "u")
    assert np.array_equal(np.array(s.timestamps), g["timestamps"])


def test_four_quadrants_64():
    g = group(TRAJ, "quad64")
    s = solver_from(g["u0"], (64, 64, 1), (3, 3, 0), (O, O, P), riemann_solver="hllc",
                    a_posteriori_slope_limiting=True)
    flags = []
    _record_flags(s, flags)
    s.ssprk3(n=3)
    bits = np.stack([np.packbits(f.ravel()) for f in flags])
    assert np.array_equal(bits, g["trouble_bits"])
    assert_bits_equal(s.u.cpu().numpy(), g["u_final"], "u")


def test_shu_osher_p5():
    g = group(TRAJ, "shuosher")
    s = solver_from(g["u0"], (512, 1, 1), (5, 0, 0), (O, P, P), x=(-5, 5), riemann_solver="hllc",
                    a_posteriori_slope_limiting=True)
    assert s.fused
    s.ssprk3(n=6)
    assert_bits_equal(s.u.cpu().numpy(), g["u_final"], "u")
    assert np.array_equal(np.array(s.timestamps), g["timestamps"])


@pytest.mark.parametrize("integ", ["euler", "ssprk2", "rk4"])
def test_other_steppers(integ):
    g = group(TRAJ, integ)
    s = solver_from(g["u0"], (64, 1, 1), (2, 0, 0), (O, P, P), riemann_solver="hllc",
                    a_posteriori_slope_limiting=True)
    getattr(s, integ)(n=4)
    assert_bits_equal(s.u.cpu().numpy(), g["u_final"], "u")
    assert np.array_equal(np.array(s.timestamps), g["timestamps"])
