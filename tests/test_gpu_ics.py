"""Initial conditions built on the GPU (fvhoe/fv.py:227-277 and the IC classes of
fvhoe/initial_conditions.py on device tensors) against the reference's host
values (tests/golden/ics.npz): bit for bit where the IC uses only IEEE
arithmetic, comparisons and sqrt; within 1e-14 relative where it calls sin / exp."""

import warnings

import numpy as np
import pytest

from cases import IC_CASES
from fvhoe_amd import initial_conditions as ics
from fvhoe_amd.boundary_conditions import BoundaryCondition
from fvhoe_amd.solver import EulerSolver
from gpu_helpers import assert_bits_equal
from helpers import load

pytestmark = pytest.mark.gpu

GOLD = load("ics.npz")
TRANSCENDENTAL = {"sinus_3d", "shu_osher_z", "kelvin_helmholtz"}


def _close(a, b, what):
    err = np.abs(a - b).max() / np.abs(b).max()
    print(f"[{what}] max rel err {err:.2e}", flush=True)
    assert err <= 1e-14, what


@pytest.mark.parametrize("case", IC_CASES, ids=[c[0] for c in IC_CASES])
def test_device_ic_matches_reference(case):
    key, cls, kw, n, p, dom = case
    ic = getattr(ics, cls)(**kw)
    X, Y, Z = ics.device_mesh(n, *dom)
    h = tuple((b - a) / k for (a, b), k in zip(dom, n))
    w0 = ic.build_ic()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        w = w0(X, Y, Z).cpu().numpy()
        u0 = ics.fv_average(lambda x, y, z: ics.primitives_to_conservatives(w0(x, y, z), 1.4), X, Y, Z, h, p)
    u0 = u0.cpu().numpy()
    if key in TRANSCENDENTAL:
        _close(w, GOLD[f"{key}/w"], key + " w")
        _close(u0, GOLD[f"{key}/u0"], key + " u0")
    else:
        assert_bits_equal(w, GOLD[f"{key}/w"], key + " w")
        assert_bits_equal(u0, GOLD[f"{key}/u0"], key + " u0")


def test_solver_ic_device():
    kw = dict(nx=32, ny=32, nz=32, px=7, py=7, pz=7, gamma=1.4, progress_bar=False)
    w0 = ics.Sinus(dims="xyz", rho_min_max=(0.5, 1.5), vx=1, vy=1, vz=1)
    host = EulerSolver(w0=w0, **kw)
    dev = EulerSolver(w0=w0, ic_device=True, **kw)
    _close(dev.u.cpu().numpy(), host.u.cpu().numpy(), "sinus 32^3 p=7")
    kw = dict(nx=64, px=3, gamma=1.4, progress_bar=False, bc=BoundaryCondition(x=("ic", "outflow")))
    host = EulerSolver(w0=ics.Shock1D(), **kw)
    dev = EulerSolver(w0=ics.Shock1D(), ic_device=True, **kw)
    assert_bits_equal(dev.u.cpu().numpy(), host.u.cpu().numpy(), "shock tube")
    host.integrator = dev.integrator = "ssprk3"
    host.take_step()
    dev.take_step()
    assert_bits_equal(dev.u.cpu().numpy(), host.u.cpu().numpy(), "shock tube after a step")
