"""MUSCL reconstruction, trouble detection and face broadcast on device arrays
(fvhoe/slope_limiting.py)."""

from __future__ import annotations

import ctypes
from typing import Dict, Optional, Tuple

import numpy as np

from fvhoe_amd import _lib

_VAR = {"rho": 0, "vx": 1, "vy": 2, "vz": 3, "P": 4, "mx": 1, "my": 2, "mz": 3, "E": 4}
_AXIS = {"x": 0, "y": 1, "z": 2}


def MUSCL_interpolations(u, axis: int, limiter: Optional[str] = "minmod"):
    """(left face, right face) of the limited p=1 reconstruction (fvhoe/slope_limiting.py:24-53)."""
    if limiter not in _lib.LIMITER:
        raise ValueError(f"Unknown slope limiter: {limiter}")
    x, was_np = _lib.as_device(u)
    nv, n0, n1, n2 = _lib.shape4(x)
    shape = [nv, n0, n1, n2]
    shape[axis] -= 2
    left, right = _lib.empty(shape), _lib.empty(shape)
    _lib.call("fvhoe_muscl", _lib.ptr(x), _lib.ptr(left), _lib.ptr(right), nv, n0, n1, n2, int(axis),
              _lib.LIMITER[limiter], _lib.stream())
    return _lib.back(left, was_np), _lib.back(right, was_np)


def _limit(limiter: int, du_left, du_right):
    t = _lib.torch()
    a, was_np = _lib.as_device(du_left)
    b, _ = _lib.as_device(du_right)
    if tuple(a.shape) != tuple(b.shape):
        a, b = (x.contiguous() for x in t.broadcast_tensors(a, b))
    out = _lib.empty(a.shape)
    _lib.call("fvhoe_slope_limit", limiter, _lib.ptr(a), _lib.ptr(b), _lib.ptr(out), out.numel(), _lib.stream())
    return _lib.back(out, was_np)


def minmod(du_left, du_right):
    """clip(dR / dL, 0, 1) dL, with dL moved away from 0 by 1e-16 (fvhoe/slope_limiting.py:56-70)."""
    return _limit(_lib.LIMITER["minmod"], du_left, du_right)


def moncen(du_left, du_right):
    """sign(dC) min(2|dL|, 2|dR|, |dC|) where dL dR >= 0, else 0; dC = (dL + dR) / 2
    (fvhoe/slope_limiting.py:73-84)."""
    return _limit(_lib.LIMITER["moncen"], du_left, du_right)


def pad_spec(PAD_bounds) -> Tuple[int, float, float]:
    """The PAD loop overwrites its indicator per key, so only the last key counts
    (fvhoe/slope_limiting.py:230-237)."""
    if not PAD_bounds:
        return -1, 0.0, 0.0
    var, (lo, hi) = list(PAD_bounds.items())[-1]
    return _VAR[var], float(lo), float(hi)


def nad_mask(NAD_vars) -> int:
    if NAD_vars is None:
        return 31
    m = 0
    for v in NAD_vars:
        if v not in _VAR:
            raise ValueError(f"Variable '{v}' not found.")
        m |= 1 << _VAR[v]
    return m


def dims_mask(dims: str) -> int:
    m = 0
    for d in dims:
        m |= 1 << _AXIS[d]
    return m


def detect_troubled_cells(u, u_candidate, dims: str = "xyz", NAD_eps: float = 1e-5, mode: str = "global",
                          range_type: str = "relative", NAD_vars: tuple = None,
                          PAD_bounds: Dict[str, Tuple[float, float]] = None, SED: bool = True,
                          SED_eps: float = 1e-10, xp: str = "numpy"):
    """NAD/PAD/SED flags (fvhoe/slope_limiting.py:87-244).

    u, u_candidate: (5, ...) primitive averages with 3 ghosts per active axis.
    Returns (trouble int64, NAD magnitude, PAD magnitude) over the interior."""
    if mode not in ("global", "local"):
        raise ValueError(f"Unknown mode: {mode}")
    if range_type not in ("relative", "absolute"):
        raise ValueError(f"Unknown range_type: {range_type}")
    if xp not in ("numpy", "cupy"):
        raise ValueError(f"Unknown xp: {xp}")
    t = _lib.torch()
    w, was_np = _lib.as_device(u)
    ws, _ = _lib.as_device(u_candidate)
    _, N0, N1, N2 = _lib.shape4(w)
    mask = dims_mask(dims)
    n = [N0 - (6 if mask & 1 else 0), N1 - (6 if mask & 2 else 0), N2 - (6 if mask & 4 else 0)]
    trouble = _lib.empty(n, dtype=t.int8)
    nad = _lib.empty(n)
    pad = _lib.empty(n)
    rng = None
    if mode == "global":
        rng = _lib.empty((10,))
        scratch = _lib.empty((10,), dtype=t.int64)
        nc = N0 * N1 * N2
        _lib.call("fvhoe_minmax", _lib.ptr(w), 5, nc, _lib.ptr(rng), _lib.ptr(scratch), _lib.stream())
    pv, plo, phi = pad_spec(PAD_bounds)
    if min(n) > 0:
        _lib.call("fvhoe_detect_troubled_cells", _lib.ptr(w), _lib.ptr(ws), N0, N1, N2, mask, _lib.ptr(rng),
                  0 if mode == "global" else 1, 0 if range_type == "relative" else 1, float(NAD_eps),
                  nad_mask(NAD_vars), pv, plo, phi, int(bool(SED)), float(SED_eps), _lib.ptr(trouble),
                  _lib.ptr(nad), _lib.ptr(pad), _lib.stream())
    trouble = trouble.to(t.int64)
    return _lib.back(trouble, was_np), _lib.back(nad, was_np), _lib.back(pad, was_np)


def _shift_max(out, s1, coef, gt, s2):
    """out = max(coef * g(s1), s2) elementwise over equal-shape 3-D views."""
    ext = (ctypes.c_int * 3)(*[int(v) for v in out.shape])
    if 0 in out.shape:
        return
    st = [(ctypes.c_longlong * 3)(*[int(v) for v in x.stride()]) for x in (out, s1, s2)]
    _lib.call("fvhoe_shift_max", _lib.ptr(out), st[0], 0, _lib.ptr(s1), st[1], 0, float(coef), int(gt),
              _lib.ptr(s2), st[2], 0, ext, _lib.stream())


def _np_assign(dst, src):
    """dst[...] = src under NumPy's assignment broadcasting (ValueError where NumPy raises one)."""
    try:
        ok = np.broadcast_shapes(tuple(src.shape), tuple(dst.shape)) == tuple(dst.shape)
    except ValueError:
        ok = False
    if not ok:
        raise ValueError(f"could not broadcast input array from shape {tuple(src.shape)} into shape "
                         f"{tuple(dst.shape)}")
    if dst.numel():
        dst.copy_(src.expand(dst.shape))


def _convex_1d(t1):
    """fvhoe/slope_limiting.py:387-407 on a (n+4, 1, 1) device array."""
    th = t1.clone()
    for dst, src, coef, gt, from_t in (
        (slice(None, -1), slice(1, None), 0.75, 0, True), (slice(1, None), slice(None, -1), 0.75, 0, True),
        (slice(None, -1), slice(1, None), 0.25, 1, False), (slice(1, None), slice(None, -1), 0.25, 1, False),
    ):
        tmp = th.clone()
        _shift_max(th[dst], (t1 if from_t else tmp)[src], coef, gt, tmp[dst])
    out = _lib.empty((th.shape[0] - 3, 1, 1))
    _shift_max(out, th[1:-2], 1.0, 0, th[2:-1])
    return out


def _convex_2d(t2):
    """fvhoe/slope_limiting.py:410-442 on a (nx+4, ny+4, 1) device array."""
    th = t2.clone()
    S = slice
    steps = [
        ((S(None), S(None, -1)), (S(None), S(1, None)), 0.75, 0, True),
        ((S(None), S(1, None)), (S(None), S(None, -1)), 0.75, 0, True),
        ((S(None, -1), S(None)), (S(1, None), S(None)), 0.75, 0, True),
        ((S(1, None), S(None)), (S(None, -1), S(None)), 0.75, 0, True),
        ((S(None, -1), S(None, -1)), (S(1, None), S(1, None)), 0.5, 0, True),
        ((S(None, -1), S(1, None)), (S(1, None), S(None, -1)), 0.5, 0, True),
        ((S(1, None), S(None, -1)), (S(None, -1), S(1, None)), 0.5, 0, True),
        ((S(1, None), S(1, None)), (S(None, -1), S(None, -1)), 0.5, 0, True),
        ((S(None), S(None, -1)), (S(None), S(1, None)), 0.25, 1, False),
        ((S(None), S(1, None)), (S(None), S(None, -1)), 0.25, 1, False),
        ((S(None, -1), S(None)), (S(1, None), S(None)), 0.25, 1, False),
        ((S(1, None), S(None)), (S(None, -1), S(None)), 0.25, 1, False),
    ]
    for dst, src, coef, gt, from_t in steps:
        tmp = th.clone()
        _shift_max(th[dst], (t2 if from_t else tmp)[src], coef, gt, tmp[dst])
    ox = _lib.empty((th.shape[0] - 3, th.shape[1] - 4, 1))
    oy = _lib.empty((th.shape[0] - 4, th.shape[1] - 3, 1))
    _shift_max(ox, th[2:-1, 2:-2], 1.0, 0, th[1:-2, 2:-2])
    _shift_max(oy, th[2:-2, 2:-1], 1.0, 0, th[2:-2, 1:-2])
    return ox, oy


def convex_1d_broadcast_to_troubled_interfaces(trouble):
    """Face weights theta in {0, 0.25, 0.75, 1} of the convex 1-D broadcast (fvhoe/slope_limiting.py:387-407):
    trouble of shape (nx + 4,) (two ghost cells per side) -> theta of shape (nx + 1,)."""
    x, was_np = _lib.as_device(trouble)
    if x.dim() != 1 or x.shape[0] < 4:
        raise ValueError(f"expected trouble of shape (nx + 4,), got {tuple(x.shape)}")
    out = _convex_1d(x.reshape(-1, 1, 1).contiguous()).reshape(-1)
    return _lib.back(out, was_np)


def convex_2d_broadcast_to_troubled_interfaces(trouble):
    """Face weights of the convex 2-D broadcast (fvhoe/slope_limiting.py:410-442): trouble of shape
    (nx + 4, ny + 4) -> (theta_x (nx + 1, ny), theta_y (nx, ny + 1))."""
    x, was_np = _lib.as_device(trouble)
    if x.dim() != 2 or min(x.shape) < 4:
        raise ValueError(f"expected trouble of shape (nx + 4, ny + 4), got {tuple(x.shape)}")
    ox, oy = _convex_2d(x.reshape(x.shape[0], x.shape[1], 1).contiguous())
    return _lib.back(ox[..., 0], was_np), _lib.back(oy[..., 0], was_np)


def broadcast_to_troubled_interfaces(trouble, dims: str = None, convex: bool = False, periodic_x: bool = False,
                                     periodic_y: bool = False, periodic_z: bool = False, xp: str = "numpy"):
    """Face weights theta (1, nx+1, ny, nz), (1, nx, ny+1, nz), (1, nx, ny, nz+1)
    (fvhoe/slope_limiting.py:247-384)."""
    t = _lib.torch()
    tr, was_np = _lib.as_device(trouble)
    nx, ny, nz = (int(s) for s in tr.shape)
    tx, ty, tz = _lib.zeros((nx + 1, ny, nz)), _lib.zeros((nx, ny + 1, nz)), _lib.zeros((nx, ny, nz + 1))
    if not convex:
        t8 = tr.to(t.int8).contiguous()
        _lib.call("fvhoe_broadcast_simple", _lib.ptr(t8), _lib.ptr(tx), _lib.ptr(ty), _lib.ptr(tz), nx, ny, nz,
                  dims_mask(dims), _lib.stream())
    else:
        xd, yd, zd = "x" in dims, "y" in dims, "z" in dims
        nd = xd + yd + zd
        if nd == 3:
            raise NotImplementedError("3D convex broadcast not implemented")
        a = _lib.zeros((nx + 4 if xd else 1, ny + 4 if yd else 1, nz + 4 if zd else 1))
        inner = tuple(slice(2, -2) if f else slice(None) for f in (xd, yd, zd))
        a[inner] = tr
        # the reference's periodic fills index all three axes as padded; on an inactive axis the target is
        # empty and NumPy's broadcasting rules decide between a no-op and a ValueError
        if periodic_x and xd:
            _np_assign(a[:2, 2:-2, 2:-2], tr[-2:])
            _np_assign(a[-2:, 2:-2, 2:-2], tr[:2])
        if periodic_y and yd:
            _np_assign(a[2:-2, :2, 2:-2], tr[:, -2:])
            _np_assign(a[2:-2, -2:, 2:-2], tr[:, :2])
        if periodic_z and zd:
            _np_assign(a[2:-2, 2:-2, :2], tr[2:-2, 2:-2, -2:])
            _np_assign(a[2:-2, 2:-2, -2:], tr[2:-2, 2:-2, :2])
        ax = [_AXIS[d] for d in dims]
        idx = [0, 0, 0]
        for k in ax:
            idx[k] = slice(None)
        sub = a[tuple(idx)]
        if nd == 1:
            out = _convex_1d(sub.reshape(-1, 1, 1).contiguous())
            tgt = {"x": tx, "y": ty, "z": tz}[dims]
            tgt.view(-1).copy_(out.view(-1))
        else:
            ox, oy = _convex_2d(sub.reshape(sub.shape[0], sub.shape[1], 1).contiguous())
            first, second = {"xy": (tx, ty), "yx": (tx, ty), "yz": (ty, tz), "zy": (ty, tz),
                             "xz": (tx, tz), "zx": (tx, tz)}[dims]
            first.view(-1).copy_(ox.view(-1))
            second.view(-1).copy_(oy.view(-1))
    return tuple(_lib.back(x.unsqueeze(0), was_np) for x in (tx, ty, tz))
