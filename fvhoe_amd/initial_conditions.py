"""Initial conditions: the reference's IC interface plus the workloads the
benchmarks and parity tests run on.

`InitialCondition` keeps the reference's abstract interface
(fvhoe/initial_conditions.py:11-81): `base_ic(*args, **kwargs)` returns
f(x, y, z) -> (5, nx, ny, nz) primitive point values, and `build_ic` appends
passive scalars.  ICs are host-side set-up that runs once; they are NumPy.

The workload builders at the bottom return conservative cell averages
directly (the reference's `conservative_ic=True, fv_ic=True` form,
fvhoe/solver.py:284-287), so seeded per-cell noise is well defined.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Callable, Dict, Tuple

import numpy as np


def primitives_to_conservatives(w: np.ndarray, gamma: float) -> np.ndarray:
    """m = rho v, E = P/(gamma-1) + 0.5 sum(m v) in the reference's operation
    order (fvhoe/hydro.py:173-196); host-side, used only to build ICs."""
    u = np.empty_like(w)
    u[0] = w[0]
    u[1:4] = w[0] * w[1:4]
    u[4] = w[4] / (gamma - 1) + 0.5 * np.sum(u[1:4] * w[1:4], axis=0)
    if w.shape[0] > 5:
        u[5:] = w[0] * w[5:]
    return u


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
            out = np.empty((5 + len(passive_scalars),) + x.shape)
            out[:5] = core(x, y, z)
            for i, name in enumerate(passive_scalars):
                out[5 + i] = passive_ic_funcs[name](x, y, z)
            return out

        return ic


class Shock1D(InitialCondition):
    """Sod-type shock tube (fvhoe/initial_conditions.py:209-250)."""

    def base_ic(self, dim="x", position=0.5, rho_left_right=(1, 0.125), P_left_right=(1, 0.1),
                v_left_right=(0, 0)):
        def f(x, y=None, z=None):
            r = {"x": x, "y": y, "z": z}[dim]
            left = r < position
            v = np.where(left, v_left_right[0], v_left_right[1])
            out = np.empty((5,) + r.shape)
            out[0] = np.where(left, rho_left_right[0], rho_left_right[1])
            out[1] = v if dim == "x" else 0
            out[2] = v if dim == "y" else 0
            out[3] = v if dim == "z" else 0
            out[4] = np.where(left, P_left_right[0], P_left_right[1])
            return out

        return f


class Sinus(InitialCondition):
    """Sinusoidal density wave (fvhoe/initial_conditions.py:84-130)."""

    def base_ic(self, dims="x", rho_min_max=(1, 2), P=1, vx=1, vy=0, vz=0):
        def f(x, y=None, z=None):
            rmin, rrange = rho_min_max[0], rho_min_max[1] - rho_min_max[0]
            r = np.zeros_like(x)
            for d, c in zip("xyz", (x, y, z)):
                r += c if d in dims else 0
            out = np.empty((5,) + r.shape)
            out[0] = rrange * (0.5 * np.sin(2 * np.pi * r) + 0.5) + rmin
            out[1] = vx
            out[2] = vy
            out[3] = vz
            out[4] = P
            return out

        return f


class ShuOsher1D(InitialCondition):
    """Shu-Osher shock / entropy-wave problem on [0, 10] along dim
    (fvhoe/initial_conditions.py:299-327): a Mach 3 shock at 1 running into
    rho = 1 + 0.2 sin(5 (x - 5))."""

    def base_ic(self, dim="x"):
        def f(x, y=None, z=None):
            r = {"x": x, "y": y, "z": z}[dim] - 5
            v = np.where(r < -4, 2.629369, 0)
            out = np.empty((5,) + r.shape)
            out[0] = np.where(r < -4, 3.857143, 1 + 0.2 * np.sin(5 * r))
            out[1] = v if dim == "x" else 0
            out[2] = v if dim == "y" else 0
            out[3] = v if dim == "z" else 0
            out[4] = np.where(r < -4, 10.33333, 1)
            return out

        return f


def uniform_mesh(n, x=(0, 1), y=(0, 1), z=(0, 1)):
    """Cell centres (fvhoe/fv.py:12-51, slab_thickness 0)."""
    mids = []
    for (a, b), k in zip((x, y, z), n):
        bnd = np.linspace(a, b, k + 1)
        mids.append(0.5 * (bnd[1:] + bnd[:-1]))
    return np.meshgrid(*mids, indexing="ij")


def fv_average(f: Callable, X, Y, Z, h, p=(0, 0, 0)) -> np.ndarray:
    """Gauss-Legendre cell averages of f with ceil((p+1)/2) points per axis
    (fvhoe/fv.py:227-277)."""
    pw = []
    for pi in p:
        pts, wts = np.polynomial.legendre.leggauss(-(-(pi + 1) // 2))
        pw.append((pts / 2, wts / 2))
    out = np.zeros_like(f(X, Y, Z))
    for xp, xw in zip(*pw[0]):
        for yp, yw in zip(*pw[1]):
            for zp, zw in zip(*pw[2]):
                out += (xw * yw * zw) * f(X + xp * h[0], Y + yp * h[1], Z + zp * h[2])
    return out


def conservative_fv_ic(ic: InitialCondition, n, p, gamma, x=(0, 1), y=(0, 1), z=(0, 1)) -> np.ndarray:
    """Conservative cell averages the reference's EulerSolver builds from a
    primitive IC (fvhoe/solver.py:292-306)."""
    X, Y, Z = uniform_mesh(n, x, y, z)
    h = tuple((b - a) / k for (a, b), k in zip((x, y, z), n))
    w0 = ic.build_ic()
    return fv_average(lambda xx, yy, zz: primitives_to_conservatives(w0(xx, yy, zz), gamma), X, Y, Z, h, p)


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
