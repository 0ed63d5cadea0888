"""Stencil sizes and weights (fvhoe/stencils.py:5-255).

The weights come from the C-ABI library's compile-time tables, the same
doubles the kernels use.
"""

from __future__ import annotations

import ctypes
from functools import lru_cache

import numpy as np

from fvhoe_amd import _lib


@lru_cache(maxsize=None)
def get_stencil_size(p: int, mode: str = "right") -> int:
    """fvhoe/stencils.py:5-22."""
    if mode == "total":
        return -2 * (-p // 2) + 1
    if mode == "right":
        return -(-p // 2)
    raise ValueError(f"{mode=}")


def _weights(p: int, kind: int) -> np.ndarray:
    out = (ctypes.c_double * 9)()
    n = ctypes.c_int(0)
    st = _lib.lib().fvhoe_stencil_weights(int(p), kind, out, ctypes.byref(n))
    if st == _lib.ERR_ARG:
        raise NotImplementedError(f"p={p}")
    _lib.check(st, "fvhoe_stencil_weights")
    return np.array(out[: n.value], dtype=np.float64)


@lru_cache(maxsize=None)
def get_conservative_interpolation_stencil_weights(p: int, pos: str) -> np.ndarray:
    """fvhoe/stencils.py:25-178."""
    if pos not in ("l", "c", "r"):
        raise ValueError(f"Invalid position '{pos}'.")
    return _weights(p, _lib.STENCIL[pos])


@lru_cache(maxsize=None)
def get_transverse_reconstruction_stencil_weights(p: int) -> np.ndarray:
    """fvhoe/stencils.py:181-255."""
    return _weights(p, _lib.STENCIL["t"])
