"""Ghost-cell fill on device arrays (fvhoe/boundary_conditions.py:154-412).

Kinds: periodic, outflow (mirror), reflective (mirror with the normal
momentum negated), free (nearest interior cell).  Slabs are filled in the
reference's x, y, z order, so corners compose the per-axis maps.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple, Union

from fvhoe_amd import _lib


@dataclass
class BoundaryCondition:
    x: Union[str, Tuple[str, str]] = "periodic"
    y: Union[str, Tuple[str, str]] = "periodic"
    z: Union[str, Tuple[str, str]] = "periodic"
    x_value: object = None
    y_value: object = None
    z_value: object = None

    def __post_init__(self):
        for d in "xyz":
            v = getattr(self, d)
            if isinstance(v, list):
                raise ValueError(f"Invalid type of {d}: list")
            if not isinstance(v, tuple):
                setattr(self, d, (v, v))
            elif len(v) != 2:
                raise ValueError(f"Invalid length of {d} tuple")
            for kind in getattr(self, d):
                if kind in ("dirichlet", "ic"):
                    raise NotImplementedError(
                        f"'{kind}' boundaries (host callables) are not on the device path")
                if kind not in _lib.BC:
                    raise ValueError(f"Unknown bc type '{kind}'")

    def codes(self):
        return [_lib.BC[k] for d in "xyz" for k in getattr(self, d)]

    def periodic(self):
        return tuple(getattr(self, d) == ("periodic", "periodic") for d in "xyz")

    def apply(self, u, gw: Tuple[int, int, int], t: float = None):
        """Pad u by gw per axis and fill the ghosts (fvhoe/boundary_conditions.py:306-402)."""
        x, was_np = _lib.as_device(u)
        nv, n0, n1, n2 = _lib.shape4(x)
        for g, n in zip(gw, (n0, n1, n2)):
            if g > n:
                raise ValueError(f"ghost width {g} exceeds the interior size {n}")
        out = _lib.empty((nv, n0 + 2 * gw[0], n1 + 2 * gw[1], n2 + 2 * gw[2]))
        codes = (_lib.ctypes.c_int * 6)(*self.codes())
        _lib.call("fvhoe_bc_apply", _lib.ptr(x), _lib.ptr(out), nv, n0, n1, n2, int(gw[0]), int(gw[1]), int(gw[2]),
                  codes, _lib.stream())
        return _lib.back(out, was_np)

    def to_dict(self) -> dict:
        return dict(x=self.x, y=self.y, z=self.z)
