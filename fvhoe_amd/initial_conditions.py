"""Initial conditions: the reference's IC interface and IC classes, plus the
workloads the benchmarks and parity tests run on.

`InitialCondition` keeps the reference's abstract interface
(fvhoe/initial_conditions.py:11-81): `base_ic(*args, **kwargs)` returns
f(x, y, z) -> (5, nx, ny, nz) primitive point values, and `build_ic` appends
passive scalars.  The IC functions, `fv_average` and
`primitives_to_conservatives` take NumPy arrays or float64 torch tensors on
the GPU alike (`xp_of` picks the array namespace), so a 256^3 or 2048^2
set-up can be built on the device (`EulerSolver(..., ic_device=True)`).  On
the host they are the reference's arithmetic bit for bit (golden vectors in
tests/golden/ics.npz); on the device every +, -, *, /, comparison and sqrt
is the same IEEE operation, and sin / exp are the device library's (within a
few ulp of NumPy's).

The workload builders at the bottom return conservative cell averages
directly (the reference's `conservative_ic=True, fv_ic=True` form,
fvhoe/solver.py:284-287), so seeded per-cell noise is well defined.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from numbers import Number
from typing import Callable, Dict, Tuple

import numpy as np


# ---------------------------------------------------------------------------
# array namespaces
# ---------------------------------------------------------------------------


class _NumpyNS:
    where = staticmethod(np.where)
    sin = staticmethod(np.sin)
    exp = staticmethod(np.exp)
    sqrt = staticmethod(np.sqrt)
    square = staticmethod(np.square)
    abs = staticmethod(np.abs)
    logical_and = staticmethod(np.logical_and)
    logical_not = staticmethod(np.logical_not)
    zeros_like = staticmethod(np.zeros_like)
    div = staticmethod(np.divide)

    @staticmethod
    def empty(shape, like):
        return np.empty(shape)

    @staticmethod
    def ones_bool(like):
        return np.ones_like(like, dtype=bool)

    @staticmethod
    def mean(a) -> float:
        return np.mean(a)


class _TorchNS:
    """The same calls on float64 GPU tensors; scalars become float64 (torch.where would make Python
    floats float32), division by a scalar divides by a device tensor (torch multiplies by the
    reciprocal of a host scalar, which is not IEEE division), and means are taken on the host with
    NumPy's summation."""

    def __init__(self, torch, device):
        self.t, self.dev = torch, device

    def _f64(self, a):
        if isinstance(a, self.t.Tensor):
            return a if a.dtype == self.t.bool else a.to(self.t.float64)
        return self.t.tensor(float(a), dtype=self.t.float64, device=self.dev)

    def where(self, c, a, b):
        return self.t.where(c, self._f64(a), self._f64(b))

    def sin(self, a):
        return self.t.sin(a)

    def exp(self, a):
        return self.t.exp(a)

    def sqrt(self, a):
        return self.t.sqrt(a)

    def square(self, a):
        return self.t.square(a)

    def abs(self, a):
        return self.t.abs(a)

    def logical_and(self, a, b):
        return self.t.logical_and(self._bool(a), self._bool(b))

    def logical_not(self, a):
        return self.t.logical_not(a)

    def _bool(self, a):
        return a if isinstance(a, self.t.Tensor) else self.t.tensor(bool(a), device=self.dev)

    def zeros_like(self, a):
        return self.t.zeros_like(a)

    def div(self, a, b):
        return a / self._f64(b)

    def empty(self, shape, like):
        return self.t.empty(tuple(shape), dtype=self.t.float64, device=self.dev)

    def ones_bool(self, like):
        return self.t.ones_like(like, dtype=self.t.bool)

    def mean(self, a) -> float:
        return np.mean(a.cpu().numpy())


def xp_of(a):
    """The array namespace of a (NumPy for anything that is not a torch tensor)."""
    if type(a).__module__.startswith("torch"):
        import torch

        return _TorchNS(torch, a.device)
    return _NumpyNS


def primitives_to_conservatives(w, gamma: float):
    """m = rho v, E = P/(gamma-1) + 0.5 sum(m v) in the reference's operation
    order (fvhoe/hydro.py:173-196; the sum over the three components runs
    left to right as NumPy's reduction over the leading axis does); set-up
    only.  Device tensors go through the library's compute_conservatives
    kernel (the same operations, IEEE division)."""
    if type(w).__module__.startswith("torch"):
        from fvhoe_amd import _lib

        wc = w.contiguous()
        u = _lib.torch().empty_like(wc)
        _lib.call("fvhoe_compute_conservatives", _lib.ptr(wc), _lib.ptr(u), int(wc.shape[0]), int(wc[0].numel()),
                  float(gamma), _lib.stream())
        return u
    u = np.empty_like(w)
    u[0] = w[0]
    u[1:4] = w[0] * w[1:4]
    mv = u[1:4] * w[1:4]
    u[4] = w[4] / (gamma - 1) + 0.5 * ((mv[0] + mv[1]) + mv[2])
    if w.shape[0] > 5:
        u[5:] = w[0] * w[5:]
    return u


# ---------------------------------------------------------------------------
# the IC interface and classes (fvhoe/initial_conditions.py)
# ---------------------------------------------------------------------------


class InitialCondition(ABC):
    """Abstract IC generator (fvhoe/initial_conditions.py:11-81)."""

    def __init__(self, *args, **kwargs):
        self.args = args
        self.kwargs = kwargs

    @abstractmethod
    def base_ic(self, *args, **kwargs) -> Callable:
        ...

    def build_ic(self, passive_scalars: Tuple[str, ...] = (), passive_ic_funcs: Dict[str, Callable] = None):
        core = self.base_ic(*self.args, **self.kwargs)
        missing = set(passive_scalars) - set((passive_ic_funcs or {}).keys())
        if missing:
            raise ValueError(f"Missing IC definitions for passive scalars: {missing}")

        def ic(x, y, z):
            out = xp_of(x).empty((5 + len(passive_scalars),) + tuple(x.shape), x)
            out[:5] = core(x, y, z)
            for i, name in enumerate(passive_scalars):
                out[5 + i] = passive_ic_funcs[name](x, y, z)
            return out

        return ic


def _fill(xp, shape, like, rho, vx, vy, vz, P):
    out = xp.empty((5,) + tuple(shape), like)
    out[0], out[1], out[2], out[3], out[4] = rho, vx, vy, vz, P
    return out


class Sinus(InitialCondition):
    """Sinusoidal density wave along the sum of the dims coordinates
    (fvhoe/initial_conditions.py:84-125)."""

    def base_ic(self, dims="x", rho_min_max=(1, 2), P=1, vx=1, vy=0, vz=0):
        def f(x, y=None, z=None):
            xp = xp_of(x)
            lo, rng = rho_min_max[0], rho_min_max[1] - rho_min_max[0]
            r = xp.zeros_like(x)
            for d, c in zip("xyz", (x, y, z)):
                r += c if d in dims else 0
            return _fill(xp, r.shape, x, rng * (0.5 * xp.sin(2 * np.pi * r) + 0.5) + lo, vx, vy, vz, P)

        return f


class Square(InitialCondition):
    """Square density wave: rho_max inside (0.25, 0.75) on every dims axis
    (fvhoe/initial_conditions.py:128-170)."""

    def base_ic(self, dims="x", rho_min_max=(1, 2), P=1, vx=1, vy=0, vz=0):
        def f(x, y=None, z=None):
            xp = xp_of(x)
            inside = xp.ones_bool(x)
            for d, r in zip("xyz", (x, y, z)):
                inside &= xp.logical_and(r > 0.25, r < 0.75) if d in dims else True
            return _fill(xp, x.shape, x, xp.where(inside, rho_min_max[1], rho_min_max[0]), vx, vy, vz, P)

        return f


class SlottedDisk(InitialCondition):
    """Zalesak's slotted disk in solid rotation about (0.5, 0.5)
    (fvhoe/initial_conditions.py:173-206)."""

    def base_ic(self, rho_min_max=(1, 2), P=1):
        def f(x, y=None, z=None):
            xp = xp_of(x)
            xc, yc = x - 0.5, y - 0.5
            rsq = xp.square(xc) + xp.square(y - 0.75)
            inside = rsq < 0.15 ** 2
            inside &= xp.logical_not(xp.logical_and(xp.abs(xc) < 0.025, y < 0.85))
            return _fill(xp, rsq.shape, x, xp.where(inside, rho_min_max[1], rho_min_max[0]), -yc, xc, 0, P)

        return f


class Shock1D(InitialCondition):
    """Sod-type shock tube along dim (fvhoe/initial_conditions.py:209-250)."""

    def base_ic(self, dim="x", position=0.5, rho_left_right=(1, 0.125), P_left_right=(1, 0.1),
                v_left_right=(0, 0)):
        def f(x, y=None, z=None):
            r = {"x": x, "y": y, "z": z}[dim]
            xp = xp_of(r)
            left = r < position
            v = xp.where(left, v_left_right[0], v_left_right[1])
            return _fill(xp, r.shape, r, xp.where(left, rho_left_right[0], rho_left_right[1]),
                         v if dim == "x" else 0, v if dim == "y" else 0, v if dim == "z" else 0,
                         xp.where(left, P_left_right[0], P_left_right[1]))

        return f


class DoubleShock1D(InitialCondition):
    """Two interfaces along dim, three constant states
    (fvhoe/initial_conditions.py:253-296)."""

    def base_ic(self, dim="x", positions=(0.1, 0.9), rhos=(1, 1, 1), vs=(0, 0, 0), Ps=(10 ** 3, 10 ** -2, 10 ** 2)):
        def f(x, y=None, z=None):
            r = {"x": x, "y": y, "z": z}[dim]
            xp = xp_of(r)

            def three(vals):
                return xp.where(r < positions[0], vals[0], xp.where(r < positions[1], vals[1], vals[2]))

            v = three(vs)
            return _fill(xp, r.shape, r, three(rhos), v if dim == "x" else 0, v if dim == "y" else 0,
                         v if dim == "z" else 0, three(Ps))

        return f


class ShuOsher1D(InitialCondition):
    """Shu-Osher shock / entropy-wave problem on [0, 10] along dim
    (fvhoe/initial_conditions.py:299-327): a Mach 3 shock at 1 running into
    rho = 1 + 0.2 sin(5 (x - 5))."""

    def base_ic(self, dim="x"):
        def f(x, y=None, z=None):
            r = {"x": x, "y": y, "z": z}[dim] - 5
            xp = xp_of(r)
            v = xp.where(r < -4, 2.629369, 0)
            return _fill(xp, r.shape, r, xp.where(r < -4, 3.857143, 1 + 0.2 * xp.sin(5 * r)),
                         v if dim == "x" else 0, v if dim == "y" else 0, v if dim == "z" else 0,
                         xp.where(r < -4, 10.33333, 1))

        return f


class KelvinHelmholtz2D(InitialCondition):
    """Shear layers at y = 0.25, 0.75 with a Gaussian-localized vy mode
    (fvhoe/initial_conditions.py:331-366)."""

    def base_ic(self, sigma=0.05 * np.sqrt(2), w0=0.1):
        def f(x, y=None, z=None):
            xp = xp_of(x)
            inner = xp.logical_and(0.25 < y, y < 0.75)
            vy = w0 * xp.sin(4 * np.pi * x) * (xp.exp(xp.div(-xp.square(y - 0.25), 2 * sigma * sigma))
                                               + xp.exp(xp.div(-xp.square(y - 0.75), 2 * sigma * sigma)))
            return _fill(xp, inner.shape, x, xp.where(inner, 2, 1), xp.where(inner, 0.5, -0.5), vy, 0, 2.5)

        return f


class DoubleMachReflection2D(InitialCondition):
    """Mach 10 shock at 60 degrees through x = 1/6 at y = 0
    (fvhoe/initial_conditions.py:369-399)."""

    def base_ic(self, gamma=1.4):
        def f(x, y=None, z=None):
            xp = xp_of(x)
            xs = x - xp.div(y, np.tan(np.pi / 3))
            post = xs < 1 / 6
            return _fill(xp, x.shape, x, xp.where(post, 8, gamma), xp.where(post, 7.145, 0),
                         xp.where(post, -8.25 / 2, 0), 0, xp.where(post, 116.5, 1))

        return f


class ShockTube(InitialCondition):
    """A ball (mode "tube") or box (mode "cube") of one state in another
    (fvhoe/initial_conditions.py:402-481)."""

    def base_ic(self, dims="x", center=0.0, radius=0.5, rho_in_out=(1, 0.125), P_in_out=(1, 0.1),
                vx_in_out=(0, 0), vy_in_out=(0, 0), vz_in_out=(0, 0), mode="tube", x_cube=None, y_cube=None,
                z_cube=None):
        def f(x, y=None, z=None):
            xp = xp_of(x)
            if mode == "tube":
                c = (center,) if isinstance(center, Number) else center
                if len(c) != len(dims):
                    raise ValueError("center must have the same length as dims")
                c3 = [np.nan, np.nan, np.nan]
                for i, d in enumerate(dims):
                    c3["xyz".index(d)] = c[i]
                xc = x - c3[0] if "x" in dims else 0
                yc = y - c3[1] if "y" in dims else 0
                zc = z - c3[2] if "z" in dims else 0
                inside = xc ** 2 + yc ** 2 + zc ** 2 <= radius ** 2
            elif mode == "cube":
                if x_cube is None and y_cube is None and z_cube is None:
                    raise ValueError("at least one cube dimension must be specified")
                inside = xp.ones_bool(x)
                for r, cube in ((x, x_cube), (y, y_cube), (z, z_cube)):
                    if cube is not None:
                        inside &= xp.logical_and(r >= cube[0], r <= cube[1])
            return _fill(xp, x.shape, x, xp.where(inside, *rho_in_out), xp.where(inside, *vx_in_out),
                         xp.where(inside, *vy_in_out), xp.where(inside, *vz_in_out), xp.where(inside, *P_in_out))

        return f


class SedovBlast(InitialCondition):
    """Energy E1 deposited in the corner cell or the central 2^d cells of the
    mesh the function is called on (fvhoe/initial_conditions.py:484-538)."""

    def base_ic(self, mode="corner", dims="xy", rho0=1.0, E0=1e-5, E1=1.0):
        def f(x, y=None, z=None):
            xp = xp_of(x)
            Nx, Ny, Nz = x.shape
            hx = xp.mean(x[1:, :, :] - x[:-1, :, :])
            hy = xp.mean(y[:, 1:, :] - y[:, :-1, :])
            hz = xp.mean(z[:, :, 1:] - z[:, :, :-1])
            Emax = E1
            Emax /= 2 * hx if "x" in dims else 1
            Emax /= 2 * hy if "y" in dims else 1
            Emax /= 2 * hz if "z" in dims else 1
            out = _fill(xp, x.shape, x, rho0, 0, 0, 0, E0)
            if mode == "corner":
                peak = (0, 0, 0)
            elif mode == "center":
                peak = tuple(slice(None) if d not in dims else slice(N // 2 - 1, N // 2 + 1)
                             for d, N in zip("xyz", (Nx, Ny, Nz)))
            out[4][peak] = Emax
            return out

        return f


class AthenaBlast(InitialCondition):
    """Pressure 10 inside r < 0.1 of the origin, 0.1 outside
    (fvhoe/initial_conditions.py:541-570)."""

    def base_ic(self, mode="corner", dims="xy", rho0=1.0, E0=1e-5, E1=1.0):
        def f(x, y=None, z=None):
            xp = xp_of(x)
            r = xp.sqrt(xp.square(x) + xp.square(y))
            return _fill(xp, r.shape, x, 1, 0, 0, 0, xp.where(r < 0.1, 10, 0.1))

        return f


# ---------------------------------------------------------------------------
# meshes and cell averages
# ---------------------------------------------------------------------------


def uniform_mesh(n, x=(0, 1), y=(0, 1), z=(0, 1)):
    """Cell centres (fvhoe/fv.py:12-51, slab_thickness 0)."""
    mids = []
    for (a, b), k in zip((x, y, z), n):
        bnd = np.linspace(a, b, k + 1)
        mids.append(0.5 * (bnd[1:] + bnd[:-1]))
    return np.meshgrid(*mids, indexing="ij")


def fv_average(f: Callable, X, Y, Z, h, p=(0, 0, 0)):
    """Gauss-Legendre cell averages of f with ceil((p+1)/2) points per axis
    (fvhoe/fv.py:227-277); X, Y, Z NumPy or device tensors."""
    pw = []
    for pi in p:
        pts, wts = np.polynomial.legendre.leggauss(-(-(pi + 1) // 2))
        pw.append((pts / 2, wts / 2))
    out = xp_of(X).zeros_like(f(x=X, y=Y, z=Z))
    for xp, xw in zip(*pw[0]):
        for yp, yw in zip(*pw[1]):
            for zp, zw in zip(*pw[2]):
                out += (xw * yw * zw) * f(x=X + xp * h[0], y=Y + yp * h[1], z=Z + zp * h[2])
    return out


def device_mesh(n, x=(0, 1), y=(0, 1), z=(0, 1)):
    """uniform_mesh as float64 device tensors (the same doubles)."""
    from fvhoe_amd import _lib

    torch = _lib.torch()
    return tuple(torch.from_numpy(np.ascontiguousarray(a)).to(_lib.device()) for a in uniform_mesh(n, x, y, z))


def conservative_fv_ic(ic: InitialCondition, n, p, gamma, x=(0, 1), y=(0, 1), z=(0, 1), device: bool = False):
    """Conservative cell averages the reference's EulerSolver builds from a
    primitive IC (fvhoe/solver.py:292-306); device=True builds them on the
    GPU and returns a device tensor."""
    X, Y, Z = device_mesh(n, x, y, z) if device else uniform_mesh(n, x, y, z)
    h = tuple((b - a) / k for (a, b), k in zip((x, y, z), n))
    w0 = ic.build_ic()
    return fv_average(lambda x, y, z: primitives_to_conservatives(w0(x, y, z), gamma), X, Y, Z, h, p)


# ---------------------------------------------------------------------------
# benchmark / parity workloads
# ---------------------------------------------------------------------------


def sod_shock_tube(n: int = 256, p: int = 3, gamma: float = 1.4) -> np.ndarray:
    """Sod tube on [0,1]: (1,0,1) | (0.125,0,0.1) at x=0.5."""
    return conservative_fv_ic(Shock1D(dim="x"), (n, 1, 1), (p, 0, 0), gamma)


def kelvin_helmholtz(n: int = 2048, seed: int = 0, gamma: float = 1.4) -> np.ndarray:
    """Periodic KH on [0,1]^2: rho=2, vx=-0.5 where |y-0.5|<0.25, else rho=1,
    vx=+0.5; P=2.5; vy = 0.01 sin(4 pi x) + U(-1e-3, 1e-3) seeded per cell.
    Point values at cell centres are taken as the cell averages."""
    X, Y, _ = uniform_mesh((n, n, 1))
    rng = np.random.Generator(np.random.PCG64(seed))
    noise = rng.uniform(-1e-3, 1e-3, size=X.shape)
    inner = np.abs(Y - 0.5) < 0.25
    w = np.empty((5,) + X.shape)
    w[0] = np.where(inner, 2.0, 1.0)
    w[1] = np.where(inner, -0.5, 0.5)
    w[2] = 0.01 * np.sin(4 * np.pi * X) + noise
    w[3] = 0.0
    w[4] = 2.5
    return primitives_to_conservatives(w, gamma)


def density_wave_3d(n: int = 256, p: int = 7, gamma: float = 1.4) -> np.ndarray:
    """rho = 1 + 0.5 sin(2 pi (x+y+z)), v = (1,1,1), P = 1 on periodic [0,1]^3,
    Gauss-Legendre cell averages (the reference's Sinus IC)."""
    return conservative_fv_ic(Sinus(dims="xyz", rho_min_max=(0.5, 1.5), P=1, vx=1, vy=1, vz=1),
                              (n, n, n), (p, p, p), gamma)


def four_quadrants(n: int = 1024, seed: int = 0, gamma: float = 1.4) -> np.ndarray:
    """Four-quadrant Riemann problem on [0,1]^2 split at (0.5, 0.5); per quadrant
    rho, P ~ U[0.1, 2] and u, v ~ U[-1, 1] from the seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rho = rng.uniform(0.1, 2.0, 4)
    P = rng.uniform(0.1, 2.0, 4)
    vx = rng.uniform(-1.0, 1.0, 4)
    vy = rng.uniform(-1.0, 1.0, 4)
    X, Y, _ = uniform_mesh((n, n, 1))
    q = (X >= 0.5).astype(np.int64) + 2 * (Y >= 0.5).astype(np.int64)
    w = np.empty((5,) + X.shape)
    w[0], w[1], w[2], w[3], w[4] = rho[q], vx[q], vy[q], 0.0, P[q]
    return primitives_to_conservatives(w, gamma)


class ShuOsherInstance(InitialCondition):
    """The Shu-Osher-type tube of the ensemble workload on [-5, 5] as a pointwise primitive IC: left
    (x < -4) (3.857143, 2.629369, 10.33333), right rho = 1 + eps sin(k x), u = 0, P = 1."""

    def base_ic(self, eps=0.2, k=5.0):
        def f(x, y=None, z=None):
            xp = xp_of(x)
            left = x < -4
            return _fill(xp, x.shape, x, xp.where(left, 3.857143, 1 + eps * xp.sin(k * x)),
                         xp.where(left, 2.629369, 0.0), 0.0, 0.0, xp.where(left, 10.33333, 1.0))

        return f


def shu_osher_instance(n: int, eps: float, k: float, gamma: float = 1.4) -> np.ndarray:
    """Shu-Osher-type tube on [-5,5]: left (x<-4) (3.857143, 2.629369, 10.33333);
    right rho = 1 + eps sin(k x), u = 0, P = 1 (cell-centre values)."""
    X, _, _ = uniform_mesh((n, 1, 1), x=(-5, 5))
    left = X < -4
    w = np.empty((5,) + X.shape)
    w[0] = np.where(left, 3.857143, 1 + eps * np.sin(k * X))
    w[1] = np.where(left, 2.629369, 0.0)
    w[2] = 0.0
    w[3] = 0.0
    w[4] = np.where(left, 10.33333, 1.0)
    return primitives_to_conservatives(w, gamma)


def shu_osher_ensemble(count: int = 512, seed: int = 0, nmin: int = 256, nmax: int = 8192):
    """Sizes N ~ U{nmin..nmax}, eps ~ U[0.1,0.3], k ~ U[3,7] per instance."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = rng.integers(nmin, nmax + 1, size=count)
    eps = rng.uniform(0.1, 0.3, size=count)
    ks = rng.uniform(3.0, 7.0, size=count)
    return [(int(n), float(e), float(k)) for n, e, k in zip(ns, eps, ks)]
