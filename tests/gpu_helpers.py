"""Helpers for the GPU parity tests."""

import numpy as np

from fvhoe_amd import initial_conditions as ics
from fvhoe_amd.boundary_conditions import BoundaryCondition
from fvhoe_amd.solver import EulerSolver


class ArrayIC(ics.InitialCondition):
    def base_ic(self, arr):
        return lambda x, y, z: arr.copy()


def solver_from(u0, n, p, bc, fused=None, **kw):
    kw = dict(kw)
    kw.setdefault("gamma", 1.4)
    return EulerSolver(w0=ArrayIC(np.asarray(u0)), nx=n[0], ny=n[1], nz=n[2], px=p[0], py=p[1], pz=p[2],
                       bc=BoundaryCondition(x=bc[0], y=bc[1], z=bc[2]), conservative_ic=True, fv_ic=True,
                       progress_bar=False, fused=fused, **kw)


def bits(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return a.view(np.int64)


def assert_bits_equal(a, b, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if not np.array_equal(bits(a), bits(b)):
        d = np.abs(a - b)
        raise AssertionError(f"{what}: {int((bits(a) != bits(b)).sum())} values differ, max |diff| {np.nanmax(d)}")
