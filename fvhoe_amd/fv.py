"""Stencil sweeps on device arrays (fvhoe/fv.py:106-224).

Each function takes a 4-D (nvars, nx, ny, nz) float64 array -- a torch
tensor on the GPU, or a NumPy array, in which case the result comes back as
NumPy -- and runs one HIP kernel per sweep.  A float32 device tensor (an
intermediate of the solver's FP32 bulk form) is swept in float32 and stays
float32.
"""

from __future__ import annotations

from functools import lru_cache
from itertools import product
from typing import Callable, Tuple

import numpy as np

from fvhoe_amd import _lib
from fvhoe_amd import initial_conditions as _ics
from fvhoe_amd.initial_conditions import uniform_mesh
from fvhoe_amd.stencils import get_stencil_size


def fv_uniform_meshgen(n, x=(0, 1), y=(0, 1), z=(0, 1), slab_thickness=(0, 0, 0)):
    """Cell-centre meshes (host NumPy, set-up only; fvhoe/fv.py:12-86).

    With slab_thickness (0, 0, 0): X, Y, Z of shape n.  Otherwise
    ((X, Y, Z), slabs) where slabs["xl"] etc. hold the centres of the ghost
    slabs: slab "xl" has slab_thickness[0] cells left of x[0] along x and the
    interior padded by the other axes' thicknesses along y and z."""
    if tuple(slab_thickness) == (0, 0, 0) and not isinstance(slab_thickness, list):
        return tuple(uniform_mesh(n, x, y, z))
    inner = tuple(uniform_mesh(n, x, y, z))
    h = ((x[1] - x[0]) / n[0], (y[1] - y[0]) / n[1], (z[1] - z[0]) / n[2])
    g = tuple(slab_thickness)
    dom = (x, y, z)
    slabs = {}
    for dim, pos in product("xyz", "lr"):
        d = "xyz".index(dim)
        cells, ext = [], []
        for e in range(3):
            lo, hi = dom[e]
            if e == d:
                cells.append(g[e])
                ext.append((lo - h[e] * g[e], lo) if pos == "l" else (hi, hi + h[e] * g[e]))
            else:
                cells.append(n[e] + 2 * g[e])
                ext.append((lo - h[e] * g[e], hi + h[e] * g[e]))
        slabs[dim + pos] = tuple(uniform_mesh(tuple(cells), *ext))
    return inner, slabs


@lru_cache(maxsize=None)
def get_symmetric_slices(ndim: int, nslices: int, axis: int) -> list:
    """The nslices shifted views along axis, left to right: cut (i, -(nslices-1)+i) for i < nslices
    (fvhoe/fv.py:88-103); index tuples for arrays of rank ndim.  The sweep kernels index the same
    windows directly."""
    from fvhoe_amd.hydro import HydroState

    if nslices < 1:
        raise ValueError(f"{nslices=}")
    hs = HydroState(ndim=ndim)
    return [hs(axis=axis, cut=(i, -(nslices - 1) + i)) for i in range(nslices)]


@lru_cache(maxsize=10)
def get_legendre_quadrature(p: int) -> Tuple[np.ndarray, np.ndarray]:
    """Gauss-Legendre points and weights on [-1/2, 1/2] with ceil((p+1)/2) points (fvhoe/fv.py:227-240)."""
    points, weights = np.polynomial.legendre.leggauss(-(-(p + 1) // 2))
    return points / 2, weights / 2


def fv_average(f: Callable, x, y, z, h, p=(0, 0, 0)):
    """Cell averages of f(x, y, z) by tensor Gauss-Legendre quadrature (fvhoe/fv.py:243-277); x, y, z are
    NumPy meshes or float64 device tensors."""
    return _ics.fv_average(f, x, y, z, h, p)


def second_order_central_difference(u, axis: int, h: float = None):
    """0.5 (u[i+1] - u[i-1]) along axis, divided by h when given (fvhoe/fv.py:280-296); u of any rank,
    the result shorter by 2 along axis."""
    x, was_np = _lib.as_device(u)
    shape = tuple(int(s) for s in x.shape)
    if not -len(shape) <= axis < len(shape):
        raise IndexError(f"axis {axis} is out of bounds for array of dimension {len(shape)}")
    axis = axis % len(shape)
    if shape[axis] < 2:
        raise ValueError(f"axis {axis} of length {shape[axis]} is shorter than the 3-point stencil")
    outer = int(np.prod(shape[:axis], dtype=np.int64))
    inner = int(np.prod(shape[axis + 1:], dtype=np.int64))
    out_shape = list(shape)
    out_shape[axis] -= 2
    out = _lib.empty(out_shape)
    _lib.call("fvhoe_central_difference", _lib.ptr(x), _lib.ptr(out), outer, shape[axis], inner,
              float(h) if h is not None else 1.0, int(h is not None), _lib.stream())
    return _lib.back(out, was_np)


def _as4(u):
    """Device tensor viewed as 4-D plus the axis shift for arrays of lower rank."""
    x, was_np = _lib.as_device_bulk(u)
    shift = 4 - x.dim()
    if shift < 0:
        raise ValueError(f"arrays of rank {x.dim()} > 4 are not supported")
    return x.reshape((1,) * shift + tuple(x.shape)), shift, was_np, tuple(x.shape)


def stencil_sweep(u, mode: str, p: int, axis: int, pos: str = None):
    """out = sum_i w_i u[i : len-(k-1)+i] along axis (fvhoe/fv.py:106-140)."""
    if mode == "interpolation":
        if pos not in ("l", "c", "r"):
            raise ValueError(f"Invalid position '{pos}'.")
        kind = _lib.STENCIL[pos]
    elif mode == "reconstruction":
        kind = _lib.STENCIL["t"]
    else:
        raise ValueError(f"{mode=}")
    if p < 0 or p > 8:
        raise NotImplementedError(f"p={p}")
    x, shift, was_np, shape = _as4(u)
    a4 = axis + shift
    k = 2 * get_stencil_size(p) + 1
    oshape = list(x.shape)
    oshape[a4] -= k - 1
    if oshape[a4] < 1:
        raise ValueError(f"axis {axis} of length {x.shape[a4]} is shorter than the {k}-point stencil")
    out = _lib.empty(oshape, dtype=x.dtype)
    _lib.call("fvhoe_stencil_sweep" + _lib.sfx(x), _lib.ptr(x), _lib.ptr(out), *x.shape, kind, p, a4, _lib.stream())
    out = out.reshape(tuple(oshape[shift:]))
    return _lib.back(out, was_np)


def conservative_interpolation(u, p: int, axis: int, pos: str = "c"):
    """fvhoe/fv.py:143-159."""
    return stencil_sweep(u, "interpolation", p, axis, pos)


def transverse_reconstruction(u, p: int, axis: int):
    """fvhoe/fv.py:162-172."""
    return stencil_sweep(u, "reconstruction", p, axis)


def _zpair(x, kind: int, p: int, axis: int, zfirst: bool):
    """Sweeps along axis (1 or 2) and along z of one kind and degree in one pass, or None where the
    pair kernel does not apply (p = 0, fewer than 64 cells along z)."""
    k = 2 * get_stencil_size(p) + 1
    oshape = list(x.shape)
    oshape[axis] -= k - 1
    oshape[3] -= k - 1
    if min(oshape) < 1:
        return None
    out = _lib.empty(oshape, dtype=x.dtype)
    name = "fvhoe_stencil_sweep_zpair" + _lib.sfx(x)
    st = getattr(_lib.lib(), name)(_lib.ptr(x), _lib.ptr(out), *x.shape, kind, p, axis, int(zfirst), _lib.stream())
    if st == _lib.ERR_UNSUPPORTED:
        return None
    _lib.check(st, name)
    return out


def _xypair(x, kind: int, p: int, first: int):
    """Sweeps along x and y (first along axis `first`) of one kind and degree in one pass, or None where
    the pair kernel does not apply (p = 0, small or very large arrays)."""
    k = 2 * get_stencil_size(p) + 1
    oshape = list(x.shape)
    oshape[1] -= k - 1
    oshape[2] -= k - 1
    if min(oshape) < 1:
        return None
    out = _lib.empty(oshape, dtype=x.dtype)
    name = "fvhoe_stencil_sweep_xypair" + _lib.sfx(x)
    st = getattr(_lib.lib(), name)(_lib.ptr(x), _lib.ptr(out), *x.shape, kind, p, first, _lib.stream())
    if st == _lib.ERR_UNSUPPORTED:
        return None
    _lib.check(st, name)
    return out


def sweep_chain(u, pos: str, steps):
    """Sweeps of one kind ("c" conservative interpolation, "t" transverse reconstruction) applied in
    order, steps = ((p, axis), ...) on a 4-D array; an x or y sweep next to a z sweep of the same degree,
    or an x sweep next to a y sweep, runs as one pass (the same values as the separate sweeps)."""
    x, was_np = _lib.as_device_bulk(u)
    kind = _lib.STENCIL[pos]
    mode = "reconstruction" if pos == "t" else "interpolation"
    steps = list(steps)
    i = 0
    while i < len(steps):
        p, a = steps[i]
        if x.dim() == 4 and i + 1 < len(steps) and steps[i + 1][0] == p and p > 0:
            a2 = steps[i + 1][1]
            pair = (a, False) if a in (1, 2) and a2 == 3 else (a2, True) if a == 3 and a2 in (1, 2) else None
            if pair is not None:
                y = _zpair(x, kind, p, *pair)
                if y is not None:
                    x = y
                    i += 2
                    continue
            z_next = i + 2 < len(steps) and steps[i + 2] == (p, 3)
            if {a, a2} == {1, 2} and not z_next:  # x and y in one pass (a following z pairs with y instead)
                y = _xypair(x, kind, p, a)
                if y is not None:
                    x = y
                    i += 2
                    continue
        x = stencil_sweep(x, mode, p, a, None if pos == "t" else pos)
        i += 1
    return _lib.back(x, was_np)


def interpolate_cell_centers(u, p=(0, 0, 0)):
    """c-sweeps x, then y, then z (fvhoe/fv.py:175-197)."""
    return sweep_chain(u, "c", ((p[0], 1), (p[1], 2), (p[2], 3)))


def interpolate_fv_averages(u, p=(0, 0, 0)):
    """transverse sweeps x, then y, then z (fvhoe/fv.py:200-224)."""
    return sweep_chain(u, "t", ((p[0], 1), (p[1], 2), (p[2], 3)))
