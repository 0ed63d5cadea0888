"""Riemann solvers on device arrays (fvhoe/riemann_solvers.py)."""

from __future__ import annotations

from fvhoe_amd import _lib

_DIM = {"x": 0, "y": 1, "z": 2}


def _solve(name, riemann_problem, gamma, dim, csq_floor):
    if dim not in _DIM:
        raise ValueError(f"dim={dim}")
    wl, was_np = _lib.as_device(riemann_problem[0])
    wr, _ = _lib.as_device(riemann_problem[1])
    if tuple(wl.shape) != tuple(wr.shape):
        raise ValueError("left and right states differ in shape")
    nc = 1
    for s in wl.shape[1:]:
        nc *= int(s)
    out = _lib.empty(wl.shape)
    _lib.call("fvhoe_riemann", _lib.ptr(wl), _lib.ptr(wr), _lib.ptr(out), wl.shape[0], nc, _lib.RIEMANN[name],
              _DIM[dim], float(gamma), float(csq_floor), _lib.stream())
    return _lib.back(out, was_np)


def advection_upwind(hs, riemann_problem, gamma: float, dim: str, csq_floor: float):
    """Upwinding without pressure (fvhoe/riemann_solvers.py:13-45)."""
    return _solve("advection_upwind", riemann_problem, gamma, dim, csq_floor)


def llf(hs, riemann_problem, gamma: float, dim: str, csq_floor: float):
    """Local Lax-Friedrichs, Toro p.331 (fvhoe/riemann_solvers.py:48-88)."""
    return _solve("llf", riemann_problem, gamma, dim, csq_floor)


def hllc(hs, riemann_problem, gamma: float, dim: str, csq_floor: float):
    """HLLC, 'David variation' (fvhoe/riemann_solvers.py:91-197)."""
    return _solve("hllc", riemann_problem, gamma, dim, csq_floor)
