"""Ghost-cell fill on device arrays (fvhoe/boundary_conditions.py:154-412).

Kinds: periodic, outflow (mirror), reflective (mirror with the normal
momentum negated), free (nearest interior cell), and the two slab kinds,
"dirichlet" (constant per-variable values, or a callable f(x, y, z, t)) and
"ic" (the solver's pointwise conservative initial condition f(x, y, z)).
Slabs are filled in the reference's x, y, z order, so a later axis decides a
corner.  Slab values are evaluated on the host, at the slab cell centres of
`slab_coords` (the reference's buffer of ghost-slab coordinates), copied to
the device once per (ghost width, time), and read there by the fill kernel.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Tuple, Union

import numpy as np

from fvhoe_amd import _lib

SLAB_KINDS = ("dirichlet", "ic")


@dataclass
class BoundaryCondition:
    x: Union[str, Tuple[str, str]] = "periodic"
    y: Union[str, Tuple[str, str]] = "periodic"
    z: Union[str, Tuple[str, str]] = "periodic"
    x_value: object = None
    y_value: object = None
    z_value: object = None
    slab_coords: dict = None
    # accepted for the reference's signature (fvhoe/boundary_conditions.py:189-190): its array manager and
    # variable map; here slabs live in device tensors and the variable order is fixed
    array_manager: object = field(default_factory=lambda: None)
    hydro_state: object = field(default_factory=lambda: None)

    def __post_init__(self):
        for d in "xyz":
            for suffix in ("", "_value"):
                v = getattr(self, d + suffix)
                if isinstance(v, list):
                    raise ValueError(f"Invalid type of {d}{suffix}: list")
                if not isinstance(v, tuple):
                    setattr(self, d + suffix, (v, v))
                elif len(v) != 2:
                    raise ValueError(f"Invalid length of {d}{suffix} tuple")
            for kind in getattr(self, d):
                if kind is not None and kind not in _lib.BC:
                    raise ValueError(f"Unknown bc type '{kind}'")
        if self.slab_coords is None:
            self.slab_buffer_size = (0, 0, 0)
        else:
            self.slab_buffer_size = (self.slab_coords["xl"][0].shape[0], self.slab_coords["yl"][0].shape[1],
                                     self.slab_coords["zl"][0].shape[2])
        self._slab_cache = {}
        self.version = 0  # bumped by reset_value: holders of device slabs fetch them again

    # ------------------------------------------------------------- settings
    def codes(self):
        """FVHOE_BC_* per side; None (no fill, the reference leaves such ghosts unset) maps to periodic,
        which only matters where a ghost width is zero."""
        return [_lib.BC[k] if k is not None else _lib.BC["periodic"] for d in "xyz" for k in getattr(self, d)]

    def periodic(self):
        return tuple(getattr(self, d) == ("periodic", "periodic") for d in "xyz")

    def has_slabs(self) -> bool:
        return any(k in SLAB_KINDS for d in "xyz" for k in getattr(self, d))

    def time_dependent(self) -> bool:
        """True if a "dirichlet" side takes a callable (evaluated at the solver's time)."""
        return any(k == "dirichlet" and callable(v) for d in "xyz"
                   for k, v in zip(getattr(self, d), getattr(self, f"{d}_value")))

    def reset_value(self, dim: str, pos: str, new_value):
        """fvhoe/boundary_conditions.py:222-238."""
        lv, rv = getattr(self, f"{dim}_value")
        if pos == "l":
            lv = new_value
        elif pos == "r":
            rv = new_value
        else:
            raise ValueError(f"Invalid pos '{pos}'")
        setattr(self, f"{dim}_value", (lv, rv))
        self._slab_cache = {}
        self.version += 1

    def trim_slabs(self, gw: Tuple[int, int, int], axis: int, pos: str):
        """Slab-centre coordinates of side (axis, pos) for ghost widths gw, cut from the
        slab_coords buffer (fvhoe/boundary_conditions.py:240-304)."""
        if self.slab_coords is None:
            raise ValueError("slab boundaries need slab_coords (fvhoe_amd.fv.fv_uniform_meshgen)")
        X, Y, Z = self.slab_coords["xyz"[axis] + pos]
        excess = [b - g for b, g in zip(self.slab_buffer_size, gw)]
        for e in range(3):
            if excess[e] < 0:
                raise ValueError(f"Cannot apply bc to region of thickness {gw[e]} with a buffer of thickness "
                                 f"{self.slab_buffer_size[e]}")
        cut = []
        for e in range(3):
            L = X.shape[e]
            if e == axis:
                cut.append(slice(L - gw[e], L) if pos == "l" else slice(0, gw[e]))
            else:
                cut.append(slice(excess[e], L - excess[e]))
        cut = tuple(cut)
        return X[cut], Y[cut], Z[cut]

    def slab_values(self, gw: Tuple[int, int, int], axis: int, pos: str, nv: int, t: float = None) -> np.ndarray:
        """Host values (nv, slab shape) of a slab side (fvhoe/boundary_conditions.py:18-57)."""
        d = "xyz"[axis]
        kind = getattr(self, d)["lr".index(pos)]
        value = getattr(self, f"{d}_value")["lr".index(pos)]
        X, Y, Z = self.trim_slabs(gw, axis, pos)
        if callable(value):
            kw = dict(x=X, y=Y, z=Z)
            if kind == "dirichlet":
                kw["t"] = t
            out = np.asarray(value(**kw), dtype=np.float64)
        elif isinstance(value, np.ndarray):
            out = np.ones((nv, *X.shape)) * value.reshape(-1, 1, 1, 1)
        else:
            raise ValueError(f"Invalid boundary of type '{value}'.")
        if out.shape != (nv, *X.shape):
            raise ValueError(f"boundary values of shape {out.shape} do not fill a slab of shape {(nv, *X.shape)}")
        return np.ascontiguousarray(out)

    def device_slabs(self, gw: Tuple[int, int, int], nv: int, t: float = None):
        """Device slab per side (None for non-slab sides and zero widths), cached per (gw, nv) and, for
        callables of "dirichlet" sides, per t.  The tensors stay alive in the cache, so kernels that read
        them asynchronously never see freed memory."""
        gw = tuple(int(g) for g in gw)
        key = (gw, nv, t if self.time_dependent() else None)
        hit = self._slab_cache.get(key)
        if hit is None:
            if self.time_dependent():  # keep one time level: drop the slabs of earlier times
                self._slab_cache = {k: v for k, v in self._slab_cache.items() if k[2] is None}
            hit = []
            for axis in range(3):
                for j, pos in enumerate("lr"):
                    kind = getattr(self, "xyz"[axis])[j]
                    if kind in SLAB_KINDS and gw[axis] > 0:
                        vals = self.slab_values(gw, axis, pos, nv, t)
                        hit.append(_lib.torch().from_numpy(vals).to(_lib.device()))
                    else:
                        hit.append(None)
            self._slab_cache[key] = hit
        return hit

    # ----------------------------------------------------------------- fill
    def apply(self, u, gw: Tuple[int, int, int], t: float = None, dtype=None):
        """Pad u by gw per axis and fill the ghosts (fvhoe/boundary_conditions.py:306-402).  dtype
        torch.float32 writes the float32 rounding of the filled state (the FP32 bulk form's first array;
        periodic / outflow / reflective / free sides only)."""
        x, was_np = _lib.as_device(u)
        f32 = dtype is not None and dtype == _lib.torch().float32
        nv, n0, n1, n2 = _lib.shape4(x)
        gw = tuple(int(g) for g in gw)
        for (d, kinds), g, n in zip(((d, getattr(self, d)) for d in "xyz"), gw, (n0, n1, n2)):
            if g == 0:
                continue
            if None in kinds:
                raise ValueError(f"ghost cells of a None boundary along {d} are undefined (ghost width {g})")
            if g > n and any(k not in SLAB_KINDS for k in kinds):
                raise ValueError(f"ghost width {g} exceeds the interior size {n}")
        out = _lib.empty((nv, n0 + 2 * gw[0], n1 + 2 * gw[1], n2 + 2 * gw[2]), dtype=dtype)
        codes = (_lib.ctypes.c_int * 6)(*self.codes())
        if f32:
            if self.has_slabs():
                raise NotImplementedError("the float32 ghost fill takes periodic / outflow / reflective / free sides")
            _lib.call("fvhoe_bc_apply_f32", _lib.ptr(x), _lib.ptr(out), nv, n0, n1, n2, *gw, codes, _lib.stream())
        elif self.has_slabs():
            slabs = self.device_slabs(gw, nv, t)
            ptrs = (_lib.ctypes.c_void_p * 6)(*[s.data_ptr() if s is not None else None for s in slabs])
            _lib.call("fvhoe_bc_apply_dirichlet", _lib.ptr(x), _lib.ptr(out), nv, n0, n1, n2, *gw, codes, ptrs,
                      _lib.stream())
        else:
            _lib.call("fvhoe_bc_apply", _lib.ptr(x), _lib.ptr(out), nv, n0, n1, n2, *gw, codes, _lib.stream())
        return _lib.back(out, was_np)

    def apply_swept(self, u, gw: Tuple[int, int, int], pos: str, p: int, axis: int, dtype=None):
        """The fill of apply() followed by one stencil sweep of kind pos ("c" or "t") and degree p along
        axis (1 or 2), in one pass (the padded array is never stored); the same values as apply() and
        fvhoe_amd.fv.stencil_sweep.  None where the fused kernel does not apply (slab sides, p = 0,
        small or very large arrays): then fill and sweep separately."""
        if self.has_slabs() or p < 1 or axis not in (1, 2):
            return None
        x, was_np = _lib.as_device(u)
        if was_np:
            return None
        nv, n0, n1, n2 = _lib.shape4(x)
        gw = tuple(int(g) for g in gw)
        for (d, kinds), g, n in zip(((d, getattr(self, d)) for d in "xyz"), gw, (n0, n1, n2)):
            if g > 0 and (None in kinds or g > n):
                return None
        shape = [nv, n0 + 2 * gw[0], n1 + 2 * gw[1], n2 + 2 * gw[2]]
        shape[axis] -= 2 * ((p + 1) // 2)
        if shape[axis] < 1:
            return None
        f32 = dtype is not None and dtype == _lib.torch().float32
        out = _lib.empty(shape, dtype=dtype)
        codes = (_lib.ctypes.c_int * 6)(*self.codes())
        name = "fvhoe_bc_stencil_sweep" + ("_f32" if f32 else "")
        st = getattr(_lib.lib(), name)(_lib.ptr(x), _lib.ptr(out), nv, n0, n1, n2, *gw, codes, _lib.STENCIL[pos], p,
                                        axis, _lib.stream())
        if st == _lib.ERR_UNSUPPORTED:
            return None
        _lib.check(st, name)
        return out

    def to_dict(self) -> dict:
        return dict(x=self.x, y=self.y, z=self.z)
