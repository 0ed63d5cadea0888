"""Smooth-extrema detector alpha on device arrays (fvhoe/smooth_extrema_detection.py)."""

from __future__ import annotations

from fvhoe_amd import _lib

_AXIS = {"x": 0, "y": 1, "z": 2}


def _sed(u, dims: str, eps: float):
    x, was_np = _lib.as_device(u)
    nv, N0, N1, N2 = _lib.shape4(x)
    mask = 0
    for d in dims:
        mask |= 1 << _AXIS[d]
    n = [N0 - (6 if mask & 1 else 0), N1 - (6 if mask & 2 else 0), N2 - (6 if mask & 4 else 0)]
    if min(n) < 0:
        raise ValueError("array shorter than the 7-point detector")
    out = _lib.empty((nv, *n))
    _lib.call("fvhoe_smooth_extrema_detector", _lib.ptr(x), _lib.ptr(out), nv, N0, N1, N2, mask, float(eps),
              _lib.stream())
    return _lib.back(out, was_np)


def compute_1d_smooth_extrema_detector(u, dim: str, eps: float = 1e-10):
    """fvhoe/smooth_extrema_detection.py:8-50 (shorter by 6 along dim)."""
    return _sed(u, dim, eps)


def compute_2d_smooth_extrema_detector(u, dims: str, eps: float = 1e-10):
    """fvhoe/smooth_extrema_detection.py:53-74 (shorter by 6 along both dims)."""
    return _sed(u, dims, eps)


def compute_3d_smooth_extrema_detector(u, eps: float = 1e-10):
    """fvhoe/smooth_extrema_detection.py:77-95."""
    return _sed(u, "xyz", eps)
