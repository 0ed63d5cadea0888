"""Hydrodynamic state layout and thermodynamics on device arrays
(fvhoe/hydro.py:7-294)."""

from __future__ import annotations

from functools import lru_cache
from typing import Tuple, Union

import numpy as np

from fvhoe_amd import _lib


class HydroState:
    """Variable -> index map and slicer (fvhoe/hydro.py:7-139): rho 0,
    v/m 1:4, P/E 4, passive scalars from 5 on."""

    def __init__(self, passive_scalars: tuple = (), ndim: int = 4):
        if len(passive_scalars) != len(set(passive_scalars)):
            raise ValueError("Duplicate passive scalar names.")
        self.passive_scalars = tuple(passive_scalars)
        self.ndim = ndim
        self.variable_map = {"rho": 0, "vx": 1, "mx": 1, "vy": 2, "my": 2, "vz": 3, "mz": 3,
                             "v": np.arange(1, 4), "m": np.arange(1, 4), "P": 4, "E": 4,
                             "active_scalars": np.arange(0, 5)}
        for i, name in enumerate(self.passive_scalars, start=5):
            self.variable_map[name] = i
        self.includes_passives = bool(self.passive_scalars)
        if self.includes_passives:
            self.variable_map["passive_scalars"] = np.arange(5, 5 + len(self.passive_scalars))
        self.nvars = 5 + len(self.passive_scalars)

    def __hash__(self):
        return id(self)

    @lru_cache(maxsize=None)
    def __call__(self, var: Union[str, Tuple[str]] = None, x=None, y=None, z=None, axis: int = None,
                 cut: Tuple[int, int] = None, step: int = None):
        sl = [slice(None)] * self.ndim
        if var is not None:
            if isinstance(var, str):
                if var not in self.variable_map:
                    raise ValueError(f"Variable '{var}' not found.")
                sl[0] = self.variable_map[var]
            elif isinstance(var, tuple):
                missing = set(var) - set(self.variable_map)
                if missing:
                    raise ValueError(f"Variables not found: {missing}")
                sl[0] = np.array([self.variable_map[v] for v in var])
            else:
                raise ValueError(f"Invalid type for var: {type(var)}")
        for i, s in zip([1, 2, 3, axis], [x, y, z, cut]):
            if s is None:
                continue
            if i >= self.ndim:
                raise ValueError(f"Invalid axis {i} for array with {self.ndim} dimensions.")
            if not isinstance(s, tuple) or len(s) != 2:
                raise ValueError(f"Expected a tuple (start, stop) for axis {i}, got {s}")
            sl[i] = slice(s[0] or None, s[1] or None, step if i == axis else None)
        return sl[0] if len(sl) == 1 else tuple(sl)


def _ncell(x):
    n = 1
    for s in x.shape[1:]:
        n *= int(s)
    return n


def compute_primitives(hs: HydroState, u, gamma: float):
    """v = m/rho, P = (gamma-1)(E - 0.5 sum(m v)) (fvhoe/hydro.py:146-170)."""
    x, was_np = _lib.as_device_bulk(u)
    out = _lib.empty(x.shape, dtype=x.dtype)
    _lib.call("fvhoe_compute_primitives" + _lib.sfx(x), _lib.ptr(x), _lib.ptr(out), x.shape[0], _ncell(x),
              float(gamma), _lib.stream())
    return _lib.back(out, was_np)


def compute_conservatives(hs: HydroState, w, gamma: float):
    """m = rho v, E = P/(gamma-1) + 0.5 sum(m v) (fvhoe/hydro.py:173-196)."""
    x, was_np = _lib.as_device(w)
    out = _lib.empty(x.shape)
    _lib.call("fvhoe_compute_conservatives", _lib.ptr(x), _lib.ptr(out), x.shape[0], _ncell(x), float(gamma),
              _lib.stream())
    return _lib.back(out, was_np)


def compute_sound_speed(w, gamma: float, csq_floor: float):
    """sqrt(max(gamma P / rho, floor)) (fvhoe/hydro.py:199-210)."""
    x, was_np = _lib.as_device(w)
    out = _lib.empty(x.shape[1:])
    _lib.call("fvhoe_compute_sound_speed", _lib.ptr(x), _lib.ptr(out), _ncell(x), float(gamma), float(csq_floor),
              _lib.stream())
    return _lib.back(out, was_np)


def compute_fluxes(hs: HydroState, w, u, gamma: float, dim: str, include_pressure: bool = True):
    """Physical flux along dim from primitives w and conservatives u (fvhoe/hydro.py:213-246):
    rho v, m v (+ P on the dim component), E v (+ v P), passives advected; gamma is not used."""
    if dim not in ("x", "y", "z"):
        raise ValueError(f"Variable 'v{dim}' not found.")
    x, was_np = _lib.as_device(w)
    y, _ = _lib.as_device(u)
    if tuple(x.shape) != tuple(y.shape) or x.shape[0] < 5:
        raise ValueError(f"w and u must share a (nvars >= 5, ...) shape, got {tuple(x.shape)}, {tuple(y.shape)}")
    if x.shape[0] != hs.nvars:
        raise ValueError(f"{x.shape[0]} variables for a HydroState of {hs.nvars}")
    out = _lib.empty(x.shape)
    _lib.call("fvhoe_compute_fluxes", _lib.ptr(x), _lib.ptr(y), _lib.ptr(out), x.shape[0], _ncell(x),
              "xyz".index(dim), int(bool(include_pressure)), _lib.stream())
    return _lib.back(out, was_np)


def advection_dt(h: float, vx: float = 0.0, vy: float = 0.0, vz: float = 0.0, CFL: float = 0.8) -> float:
    """Time step of an advection problem at Courant number CFL (fvhoe/hydro.py:249-267): scalar host
    arithmetic, CFL h / (vx + vy + vz)."""
    return CFL * h / (vx + vy + vz)


def hydro_dt_device(w, h: float, ndim: int, CFL: float, gamma: float, csq_floor: float, out=None):
    """hydro_dt with the result left on the device (1-element tensor; into out when given)."""
    x, _ = _lib.as_device_bulk(w)
    t = _lib.torch()
    dt = _lib.empty((1,)) if out is None else out
    scratch = _lib.empty((1,), dtype=t.int64)
    _lib.call("fvhoe_hydro_dt" + _lib.sfx(x), _lib.ptr(x), _ncell(x), float(h), int(ndim), float(CFL), float(gamma),
              float(csq_floor), _lib.ptr(dt), _lib.ptr(scratch), _lib.stream())
    return dt


def hydro_dt(w, h: float, ndim: int, CFL: float, gamma: float, csq_floor: float) -> float:
    """CFL h / max(|vx|+|vy|+|vz| + ndim c) (fvhoe/hydro.py:270-294)."""
    return float(hydro_dt_device(w, h, ndim, CFL, gamma, csq_floor).item())
