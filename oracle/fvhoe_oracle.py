"""NumPy restatement of FVHOE's per-RK-stage update (CPU oracle; test infrastructure).

Every function follows the reference's floating-point operation order so
that FP64 results can be compared bit for bit.  Arrays are 4-D
``(nvars, nx, ny, nz)`` C-order float64 as in the reference.  Variable order:
rho, vx|mx, vy|my, vz|mz, P|E, passives (fvhoe/hydro.py:36-63).

The FP32 bulk form (``Scheme(bulk="fp32")``, no reference counterpart) runs
the same functions on float32 intermediates: the sweeps and Riemann solves in
float32 (weights and gamma rounded to float32 once), the pressure recovery,
CFL maximum and divergence in float64 on the float32 values, the state and
the RK combine in float64.

Only tests, smoke() and bench.py's cpu_baseline leg use this module.
"""

from __future__ import annotations

import dataclasses
from fractions import Fraction
from functools import lru_cache
from typing import Dict, Optional, Tuple

import numpy as np

AXES = {"x": 1, "y": 2, "z": 3}

# ---------------------------------------------------------------------------
# stencil weights (fvhoe/stencils.py:5-255), derived from moment conditions
# ---------------------------------------------------------------------------


def stencil_reach(p: int) -> int:
    """ceil(p/2): cells on each side of a stencil (fvhoe/stencils.py:5-22)."""
    return -(-p // 2)


def _moment_weights(offsets, x0) -> list:
    """Exact weights w_k on unit cells centred at `offsets` such that
    sum_k w_k * avg_k(q) == q(x0) for every polynomial q of degree < len(offsets)."""
    n = len(offsets)
    # A[m][k] = average of x^m over cell k = ((o+1/2)^(m+1) - (o-1/2)^(m+1)) / (m+1)
    A = []
    for m in range(n):
        row = []
        for o in offsets:
            o = Fraction(o)
            hi, lo = o + Fraction(1, 2), o - Fraction(1, 2)
            row.append((hi ** (m + 1) - lo ** (m + 1)) / (m + 1))
        A.append(row)
    b = [Fraction(x0) ** m for m in range(n)]
    # Gaussian elimination over the rationals
    M = [A[i][:] + [b[i]] for i in range(n)]
    for c in range(n):
        piv = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [a - f * e for a, e in zip(M[r], M[c])]
    return [M[i][n] / M[i][i] for i in range(n)]


def _point_from_averages(p: int, x0: Fraction) -> list:
    """Rational weights on cells -c..c for the point value at x0 (even p: the
    unique symmetric degree-p stencil; odd p: mean of the two degree-p stencils)."""
    c = stencil_reach(p)
    if p % 2 == 0:
        w = _moment_weights(list(range(-c, c + 1)), x0)
    else:
        wl = _moment_weights(list(range(-c, c)), x0) + [Fraction(0)]
        wr = [Fraction(0)] + _moment_weights(list(range(-c + 1, c + 1)), x0)
        w = [(a + b) / 2 for a, b in zip(wl, wr)]
    return w


def _average_from_points(p: int) -> list:
    """Rational weights taking centre point values to a cell average.

    Even p: the unique symmetric stencil with sum_k w_k q(k) == avg_0(q) for
    deg q <= p; odd p: the p-1 stencil padded with zeros."""
    c = stencil_reach(p)
    q = p if p % 2 == 0 else p - 1
    cq = q // 2
    offs = list(range(-cq, cq + 1))
    n = len(offs)
    # solve V w = mom, V[m][k] = offs_k^m, mom[m] = avg of x^m over [-1/2, 1/2]
    M = []
    for m in range(n):
        row = [Fraction(o) ** m for o in offs]
        mom = (Fraction(1, 2) ** (m + 1) - Fraction(-1, 2) ** (m + 1)) / (m + 1)
        M.append(row + [mom])
    for col in range(n):
        piv = next(r for r in range(col, n) if M[r][col] != 0)
        M[col], M[piv] = M[piv], M[col]
        for r in range(n):
            if r != col and M[r][col] != 0:
                f = M[r][col] / M[col][col]
                M[r] = [a - f * e for a, e in zip(M[r], M[col])]
    w = [M[i][n] / M[i][i] for i in range(n)]
    pad = [Fraction(0)] * (c - cq)
    return pad + w + pad


@lru_cache(maxsize=None)
def interpolation_weights_exact(p: int, pos: str) -> tuple:
    """Exact rationals behind get_conservative_interpolation_stencil_weights
    (fvhoe/stencils.py:25-178); "r" is the reversed "l" (:42-43)."""
    if pos not in ("l", "c", "r"):
        raise ValueError(f"Invalid position '{pos}'.")
    if p < 0 or p > 8:
        raise NotImplementedError(f"p={p}")
    if p == 0:
        return (Fraction(1),)
    if pos == "r":
        return tuple(reversed(interpolation_weights_exact(p, "l")))
    if pos == "l":
        return tuple(_point_from_averages(p, Fraction(-1, 2)))
    # centre: odd p uses the p-1 stencil padded with zeros
    c = stencil_reach(p)
    if p % 2 == 0:
        return tuple(_point_from_averages(p, Fraction(0)))
    inner = _point_from_averages(p - 1, Fraction(0))
    pad = [Fraction(0)] * (c - stencil_reach(p - 1))
    return tuple(pad + inner + pad)


@lru_cache(maxsize=None)
def transverse_weights_exact(p: int) -> tuple:
    """Exact rationals behind get_transverse_reconstruction_stencil_weights
    (fvhoe/stencils.py:181-255)."""
    if p < 0 or p > 8:
        raise NotImplementedError(f"p={p}")
    if p == 0:
        return (Fraction(1),)
    return tuple(_average_from_points(p))


def interpolation_weights(p: int, pos: str) -> np.ndarray:
    """float64 weights; each entry is the correctly rounded rational, which is
    what the reference's `a / b` literals evaluate to (fvhoe/stencils.py:45-176)."""
    return np.array([float(f) for f in interpolation_weights_exact(p, pos)], dtype=np.float64)


def transverse_weights(p: int) -> np.ndarray:
    return np.array([float(f) for f in transverse_weights_exact(p)], dtype=np.float64)


# ---------------------------------------------------------------------------
# stencil sweeps (fvhoe/fv.py:88-224,280-296)
# ---------------------------------------------------------------------------


def _shift(u: np.ndarray, axis: int, start: int, stop_from_end: int):
    """u sliced along `axis` as [start : len - stop_from_end]."""
    sl = [slice(None)] * u.ndim
    sl[axis] = slice(start, u.shape[axis] - stop_from_end)
    return u[tuple(sl)]


def stencil_sweep(u: np.ndarray, weights: np.ndarray, axis: int) -> np.ndarray:
    """out = sum_i w_i * u[i : len-(k-1)+i], accumulated left to right from
    zeros (fvhoe/fv.py:106-140)."""
    k = len(weights)
    out = np.zeros_like(_shift(u, axis, 0, k - 1))
    if u.dtype == np.float32:  # float32 sums: each product and partial sum rounded to float32
        weights = np.asarray(weights, dtype=np.float64).astype(np.float32)
    for i, w in enumerate(weights):
        np.add(out, w * _shift(u, axis, i, k - 1 - i), out=out)
    return out


def conservative_interpolation(u: np.ndarray, p: int, axis: int, pos: str = "c") -> np.ndarray:
    """fvhoe/fv.py:143-159."""
    return stencil_sweep(u, interpolation_weights(p, pos), axis)


def transverse_reconstruction(u: np.ndarray, p: int, axis: int) -> np.ndarray:
    """fvhoe/fv.py:162-172."""
    return stencil_sweep(u, transverse_weights(p), axis)


def interpolate_cell_centers(u: np.ndarray, p=(0, 0, 0)) -> np.ndarray:
    """c-sweeps along x, then y, then z (fvhoe/fv.py:175-197)."""
    out = conservative_interpolation(u, p[0], 1, "c")
    out = conservative_interpolation(out, p[1], 2, "c")
    return conservative_interpolation(out, p[2], 3, "c")


def interpolate_fv_averages(u: np.ndarray, p=(0, 0, 0)) -> np.ndarray:
    """transverse sweeps along x, then y, then z (fvhoe/fv.py:200-224)."""
    out = transverse_reconstruction(u, p[0], 1)
    out = transverse_reconstruction(out, p[1], 2)
    return transverse_reconstruction(out, p[2], 3)


def second_order_central_difference(u: np.ndarray, axis: int, h: Optional[float] = None) -> np.ndarray:
    """0.5 * (u[2:] - u[:-2]) along axis, then / h when h is given (fvhoe/fv.py:280-296)."""
    out = 0.5 * (_shift(u, axis, 2, 0) - _shift(u, axis, 0, 2))
    if h is not None:
        out = out / h
    return out


# ---------------------------------------------------------------------------
# thermodynamics (fvhoe/hydro.py:146-294)
# ---------------------------------------------------------------------------


def compute_primitives(u: np.ndarray, gamma: float) -> np.ndarray:
    """fvhoe/hydro.py:146-170: v = m/rho, P = (gamma-1)(E - 0.5 sum(m v)).  A float32 u is recovered in
    float64 and rounded back to float32."""
    if u.dtype == np.float32:
        return compute_primitives(u.astype(np.float64), gamma).astype(np.float32)
    w = np.empty_like(u)
    rho = u[0]
    w[0] = rho
    w[1:4] = u[1:4] / rho
    msum = np.sum(u[1:4] * w[1:4], axis=0)
    w[4] = (gamma - 1) * (u[4] - 0.5 * msum)
    if u.shape[0] > 5:
        w[5:] = u[5:] / rho
    return w


def compute_conservatives(w: np.ndarray, gamma: float) -> np.ndarray:
    """fvhoe/hydro.py:173-196."""
    u = np.empty_like(w)
    rho = w[0]
    u[0] = rho
    u[1:4] = rho * w[1:4]
    u[4] = w[4] / (gamma - 1) + 0.5 * np.sum(u[1:4] * w[1:4], axis=0)
    if w.shape[0] > 5:
        u[5:] = rho * w[5:]
    return u


def compute_sound_speed(w: np.ndarray, gamma: float, csq_floor: float) -> np.ndarray:
    """fvhoe/hydro.py:199-210."""
    csq = gamma * w[4] / w[0]
    return np.sqrt(np.where(csq > csq_floor, csq, csq_floor))


def compute_fluxes(w: np.ndarray, u: np.ndarray, dim: str, include_pressure: bool = True) -> np.ndarray:
    """Physical Euler flux along `dim` (fvhoe/hydro.py:213-246)."""
    d = AXES[dim]
    out = np.empty_like(w)
    v = w[d]
    out[0] = v * w[0]
    out[1:4] = v * u[1:4]
    out[4] = v * u[4]
    if include_pressure:
        out[d] = out[d] + w[4]
        out[4] = out[4] + v * w[4]
    if w.shape[0] > 5:
        out[5:] = v * u[5:]
    return out


def hydro_dt(w: np.ndarray, h: float, ndim: int, CFL: float, gamma: float, csq_floor: float) -> float:
    """CFL * h / max(|vx|+|vy|+|vz| + ndim*c) over every cell given (fvhoe/hydro.py:270-294)."""
    w = w.astype(np.float64, copy=False)
    c = compute_sound_speed(w, gamma, csq_floor)
    va = np.abs(w[1:4])
    return float(CFL * h / np.max(np.sum(va, axis=0) + ndim * c))


# ---------------------------------------------------------------------------
# Riemann solvers (fvhoe/riemann_solvers.py)
# ---------------------------------------------------------------------------


def advection_upwind(wl, wr, gamma, dim, csq_floor):
    """fvhoe/riemann_solvers.py:13-45."""
    ul = compute_conservatives(wl, gamma)
    ur = compute_conservatives(wr, gamma)
    Fl = compute_fluxes(wl, ul, dim, include_pressure=False)
    Fr = compute_fluxes(wr, ur, dim, include_pressure=False)
    v = wl[AXES[dim]][np.newaxis]
    return np.where(v > 0, Fl, np.where(v < 0, Fr, 0))


def llf(wl, wr, gamma, dim, csq_floor):
    """Local Lax-Friedrichs (Toro p.331; fvhoe/riemann_solvers.py:48-88)."""
    d = AXES[dim]
    ul = compute_conservatives(wl, gamma)
    ur = compute_conservatives(wr, gamma)
    Fl = compute_fluxes(wl, ul, dim)
    Fr = compute_fluxes(wr, ur, dim)
    sl = np.abs(wl[d]) + compute_sound_speed(wl, gamma, csq_floor)
    sr = np.abs(wr[d]) + compute_sound_speed(wr, gamma, csq_floor)
    smax = np.maximum(sl, sr)
    return 0.5 * (Fl + Fr) - 0.5 * smax * (ur - ul)


def hllc(wl, wr, gamma, dim, csq_floor):
    """HLLC, 'David variation' (fvhoe/riemann_solvers.py:91-197)."""
    d = AXES[dim]
    ul = compute_conservatives(wl, gamma)
    ur = compute_conservatives(wr, gamma)
    cl = compute_sound_speed(wl, gamma, csq_floor)
    cr = compute_sound_speed(wr, gamma, csq_floor)
    cmax = np.maximum(cl, cr)
    vl, vr = wl[d], wr[d]
    rhol, rhor = wl[0], wr[0]
    Pl, Pr = wl[4], wr[4]
    El, Er = ul[4], ur[4]
    Sl = np.minimum(vl, vr) - cmax
    Sr = np.maximum(vl, vr) + cmax
    rc_L = rhol * (vl - Sl)
    rc_R = rhor * (Sr - vr)
    v_star = (rc_L * vl + rc_R * vr + Pl - Pr) / (rc_L + rc_R)
    P_star = (rc_L * Pr + rc_R * Pl + rc_L * rc_R * (vl - vr)) / (rc_L + rc_R)
    with np.errstate(divide="ignore", invalid="ignore"):
        r_starL = rhol * (Sl - vl) / (Sl - v_star)
        E_starL = ((Sl - vl) * El - Pl * vl + P_star * v_star) / (Sl - v_star)
        r_starR = rhor * (Sr - vr) / (Sr - v_star)
        E_starR = ((Sr - vr) * Er - Pr * vr + P_star * v_star) / (Sr - v_star)
    r_g = np.where(Sl > 0, rhol, np.where(v_star > 0, r_starL, np.where(Sr > 0, r_starR, rhor)))
    v_g = np.where(Sl > 0, vl, np.where(v_star > 0, v_star, np.where(Sr > 0, v_star, vr)))
    P_g = np.where(Sl > 0, Pl, np.where(v_star > 0, P_star, np.where(Sr > 0, P_star, Pr)))
    E_g = np.where(Sl > 0, El, np.where(v_star > 0, E_starL, np.where(Sr > 0, E_starR, Er)))
    out = np.empty_like(ul)
    out[0] = r_g * v_g
    for k in (1, 2, 3):
        if k == d:
            out[k] = r_g * v_g * v_g + P_g
        else:
            out[k] = r_g * v_g * np.where(v_g > 0, wl[k], np.where(v_g < 0, wr[k], 0))
    out[4] = v_g * (E_g + P_g)
    if wl.shape[0] > 5:
        out[5:] = r_g * v_g * np.where(v_g > 0, wl[5:], np.where(v_g < 0, wr[5:], 0))
    return out


RIEMANN = {"advection_upwind": advection_upwind, "llf": llf, "hllc": hllc}


# ---------------------------------------------------------------------------
# slope limiting and trouble detection (fvhoe/slope_limiting.py)
# ---------------------------------------------------------------------------


def minmod(dl: np.ndarray, dr: np.ndarray) -> np.ndarray:
    """Ratio form (fvhoe/slope_limiting.py:56-70)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = dr / np.where(dl > 0, np.where(dl > 1e-16, dl, 1e-16), np.where(dl < -1e-16, dl, -1e-16))
    ratio = np.where(ratio < 1, ratio, 1)
    return np.where(ratio > 0, ratio, 0) * dl


def moncen(dl: np.ndarray, dr: np.ndarray) -> np.ndarray:
    """fvhoe/slope_limiting.py:73-84."""
    dc = 0.5 * (dl + dr)
    slope = np.minimum(np.abs(2 * dl), np.abs(2 * dr))
    slope = np.sign(dc) * np.minimum(slope, np.abs(dc))
    return np.where(dl * dr >= 0, slope, 0)


def MUSCL_interpolations(u: np.ndarray, axis: int, limiter: Optional[str] = "minmod"):
    """(left face, right face) of the limited p=1 reconstruction (fvhoe/slope_limiting.py:24-53)."""
    u0, u1, u2 = _shift(u, axis, 0, 2), _shift(u, axis, 1, 1), _shift(u, axis, 2, 0)
    dl = u1 - u0
    dr = u2 - u1
    if limiter == "minmod":
        s = minmod(dl, dr)
    elif limiter == "moncen":
        s = moncen(dl, dr)
    elif limiter is None:
        s = 0.5 * (dl + dr)
    else:
        raise ValueError(f"Unknown slope limiter: {limiter}")
    return u1 - 0.5 * s, u1 + 0.5 * s


def compute_1d_smooth_extrema_detector(u: np.ndarray, dim: str, eps: float = 1e-10) -> np.ndarray:
    """fvhoe/smooth_extrema_detection.py:8-50; output shorter by 6 along dim."""
    axis = AXES[dim]
    dU = second_order_central_difference(u, axis)
    d2U = second_order_central_difference(dU, axis)
    dv = 0.5 * d2U
    dv = np.where(np.abs(dv) <= eps, np.where(dv >= 0, eps, -eps), dv)
    vL = _shift(dU, axis, 0, 2) - _shift(dU, axis, 1, 1)
    alphaL = -np.where(dv < 0, np.where(vL > 0, vL, 0), np.where(vL < 0, vL, 0)) / dv
    alphaL = np.where(alphaL < 1, alphaL, 1)
    vR = _shift(dU, axis, 2, 0) - _shift(dU, axis, 1, 1)
    alphaR = np.where(dv > 0, np.where(vR > 0, vR, 0), np.where(vR < 0, vR, 0)) / dv
    alphaR = np.where(np.abs(dv) <= eps, 1, alphaR)
    alphaR = np.where(alphaR < alphaL, alphaR, alphaL)
    alpha = np.minimum(alphaL, alphaR)
    return np.minimum(
        np.minimum(_shift(alpha, axis, 2, 0), _shift(alpha, axis, 1, 1)), _shift(alpha, axis, 0, 2)
    )


def _trim(u: np.ndarray, axis: int, n: int) -> np.ndarray:
    return _shift(u, axis, n, n)


def compute_2d_smooth_extrema_detector(u: np.ndarray, dims: str, eps: float = 1e-10) -> np.ndarray:
    """fvhoe/smooth_extrema_detection.py:53-74."""
    a1, a2 = AXES[dims[0]], AXES[dims[1]]
    alpha1 = compute_1d_smooth_extrema_detector(_trim(u, a2, 3), dims[0], eps)
    alpha2 = compute_1d_smooth_extrema_detector(_trim(u, a1, 3), dims[1], eps)
    return np.minimum(alpha1, alpha2)


def compute_3d_smooth_extrema_detector(u: np.ndarray, eps: float = 1e-10) -> np.ndarray:
    """fvhoe/smooth_extrema_detection.py:77-95."""
    alpha_xy = compute_2d_smooth_extrema_detector(_trim(u, 3, 3), "xy", eps)
    alpha_z = compute_1d_smooth_extrema_detector(_trim(_trim(u, 1, 3), 2, 3), "z", eps)
    return np.minimum(alpha_xy, alpha_z)


VAR_INDEX = {"rho": 0, "vx": 1, "vy": 2, "vz": 3, "P": 4, "mx": 1, "my": 2, "mz": 3, "E": 4}


def neighbour_max_min(u: np.ndarray, dims: str):
    """3^d neighbourhood max and min at the interior cells [3:-3] of every active
    axis.  Equals scipy.ndimage.maximum/minimum_filter(footprint 3^d, cval=0)
    followed by the interior slice (fvhoe/slope_limiting.py:146-162): the
    interior cells' footprints never reach the padding."""
    sl = [slice(None)] * 4
    for d in dims:
        sl[AXES[d]] = slice(3, -3)
    M = None
    m = None
    offs = [(-1, 0, 1) if (a in [AXES[d] for d in dims]) else (0,) for a in (1, 2, 3)]
    for ox in offs[0]:
        for oy in offs[1]:
            for oz in offs[2]:
                s = list(sl)
                for a, o in zip((1, 2, 3), (ox, oy, oz)):
                    if sl[a] != slice(None):
                        s[a] = slice(3 + o, u.shape[a] - 3 + o)
                v = u[tuple(s)]
                M = v.copy() if M is None else np.maximum(M, v)
                m = v.copy() if m is None else np.minimum(m, v)
    return M, m


def detect_troubled_cells(
    u: np.ndarray,
    u_candidate: np.ndarray,
    dims: str = "xyz",
    NAD_eps: float = 1e-5,
    mode: str = "global",
    range_type: str = "relative",
    NAD_vars: Optional[tuple] = None,
    PAD_bounds: Optional[Dict[str, Tuple[float, float]]] = None,
    SED: bool = True,
    SED_eps: float = 1e-10,
):
    """NAD/PAD/SED troubled-cell flags (fvhoe/slope_limiting.py:87-244).

    u, u_candidate: (5, ...) primitive fv averages with 3 ghosts per active axis.
    Returns (trouble int64, NAD magnitude, PAD magnitude) on the interior."""
    interior = [slice(None)] * 4
    for d in dims:
        interior[AXES[d]] = slice(3, -3)
    interior = tuple(interior)
    uc_in = u_candidate[interior]
    M, m = neighbour_max_min(u, dims)
    if range_type == "relative":
        if mode == "global":
            u_range = np.max(u, axis=(1, 2, 3), keepdims=True) - np.min(u, axis=(1, 2, 3), keepdims=True)
        elif mode == "local":
            u_range = M - m
        else:
            raise ValueError(f"Unknown mode: {mode}")
        ind = uc_in - m
        ind = np.minimum(ind, M - uc_in)
        nad_pv = ind < -NAD_eps * u_range
    elif range_type == "absolute":
        if mode == "global":
            lo = (1 - NAD_eps) * np.min(u, axis=(1, 2, 3), keepdims=True)
            hi = (1 + NAD_eps) * np.max(u, axis=(1, 2, 3), keepdims=True)
        elif mode == "local":
            lo = (1 - NAD_eps) * m
            hi = (1 + NAD_eps) * M
        else:
            raise ValueError(f"Unknown mode: {mode}")
        ind = np.minimum(uc_in - lo, hi - uc_in)
        nad_pv = ind < 0.0
    else:
        raise ValueError(f"Unknown range_type: {range_type}")
    if SED:
        if len(dims) == 1:
            alpha = compute_1d_smooth_extrema_detector(u_candidate, dims, SED_eps)
        elif len(dims) == 2:
            alpha = compute_2d_smooth_extrema_detector(u_candidate, dims, SED_eps)
        elif len(dims) == 3:
            alpha = compute_3d_smooth_extrema_detector(u_candidate, SED_eps)
        else:
            raise ValueError(f"Invalid dims specified: {dims}")
        nad_pv = np.where(alpha < 1, nad_pv, False)
    if NAD_vars is not None:
        keep = np.zeros((5, 1, 1, 1), dtype=bool)
        for v in NAD_vars:
            keep[VAR_INDEX[v]] = True
        nad_pv = np.where(keep, nad_pv, False)
    nad = np.any(nad_pv, axis=0)
    nad_mag = np.where(nad, -np.min(ind, axis=0), 0)
    pad_ind = np.zeros_like(nad_mag)
    if PAD_bounds is not None:
        for var, (lo_b, hi_b) in PAD_bounds.items():  # last key wins
            pad_ind = uc_in[VAR_INDEX[var]] - lo_b
            pad_ind = np.minimum(pad_ind, hi_b - uc_in[VAR_INDEX[var]])
    pad = pad_ind < 0.0
    pad_mag = np.where(pad, -pad_ind, 0)
    trouble = np.where(pad, 1, nad)
    return trouble, nad_mag, pad_mag


def broadcast_to_troubled_interfaces(trouble: np.ndarray, dims: str, convex: bool = False,
                                     periodic=(False, False, False)):
    """Face weights theta from cell flags (fvhoe/slope_limiting.py:247-442).
    Returns (theta_x (1,nx+1,ny,nz), theta_y, theta_z) float64."""
    nx, ny, nz = trouble.shape
    tx = np.zeros((nx + 1, ny, nz))
    ty = np.zeros((nx, ny + 1, nz))
    tz = np.zeros((nx, ny, nz + 1))
    if convex:
        xd, yd, zd = "x" in dims, "y" in dims, "z" in dims
        nd = xd + yd + zd
        if nd == 3:
            raise NotImplementedError("3D convex broadcast not implemented")
        a = np.zeros((nx + 4 if xd else 1, ny + 4 if yd else 1, nz + 4 if zd else 1))
        inner = tuple(slice(2, -2) if f else slice(None) for f in (xd, yd, zd))
        a[inner] = trouble
        if periodic[0] and xd:
            a[:2, 2:-2, 2:-2] = trouble[-2:, :, :]
            a[-2:, 2:-2, 2:-2] = trouble[:2, :, :]
        if periodic[1] and yd:
            a[2:-2, :2, 2:-2] = trouble[:, -2:, :]
            a[2:-2, -2:, 2:-2] = trouble[:, :2, :]
        if periodic[2] and zd:
            a[2:-2, 2:-2, :2] = trouble[2:-2, 2:-2, -2:]
            a[2:-2, 2:-2, -2:] = trouble[2:-2, 2:-2, :2]
        idx = [0, 0, 0]
        if nd == 1:
            idx[AXES[dims] - 1] = slice(None)
            t1 = _convex_1d(a[tuple(idx)])
            {"x": tx, "y": ty, "z": tz}[dims][...] = t1.reshape(
                {"x": (-1, 1, 1), "y": (1, -1, 1), "z": (1, 1, -1)}[dims])
        elif nd == 2:
            idx[AXES[dims[0]] - 1] = slice(None)
            idx[AXES[dims[1]] - 1] = slice(None)
            t1, t2 = _convex_2d(a[tuple(idx)])
            if dims in ("xy", "yx"):
                tx[...] = t1.reshape(nx + 1, ny, 1)
                ty[...] = t2.reshape(nx, ny + 1, 1)
            elif dims in ("yz", "zy"):
                ty[...] = t1.reshape(1, ny + 1, nz)
                tz[...] = t2.reshape(1, ny, nz + 1)
            elif dims in ("xz", "zx"):
                tx[...] = t1.reshape(nx + 1, 1, nz)
                tz[...] = t2.reshape(nx, 1, nz + 1)
    else:
        if "x" in dims:
            tx[1:] = trouble
            tx[:-1] = np.maximum(tx[:-1], trouble)
        if "y" in dims:
            ty[:, 1:] = trouble
            ty[:, :-1] = np.maximum(ty[:, :-1], trouble)
        if "z" in dims:
            tz[:, :, 1:] = trouble
            tz[:, :, :-1] = np.maximum(tz[:, :, :-1], trouble)
    return tx[np.newaxis], ty[np.newaxis], tz[np.newaxis]


def _convex_1d(t: np.ndarray) -> np.ndarray:
    """fvhoe/slope_limiting.py:387-407."""
    th = t.copy()
    th[:-1] = np.maximum(0.75 * t[1:], th[:-1])
    th[1:] = np.maximum(0.75 * t[:-1], th[1:])
    th[:-1] = np.maximum(0.25 * (th[1:] > 0), th[:-1])
    th[1:] = np.maximum(0.25 * (th[:-1] > 0), th[1:])
    return np.maximum(th[1:-2], th[2:-1])


def _convex_2d(t: np.ndarray):
    """fvhoe/slope_limiting.py:410-442 (in-place update order kept)."""
    th = t.copy()
    th[:, :-1] = np.maximum(0.75 * t[:, 1:], th[:, :-1])
    th[:, 1:] = np.maximum(0.75 * t[:, :-1], th[:, 1:])
    th[:-1, :] = np.maximum(0.75 * t[1:, :], th[:-1, :])
    th[1:, :] = np.maximum(0.75 * t[:-1, :], th[1:, :])
    th[:-1, :-1] = np.maximum(0.5 * t[1:, 1:], th[:-1, :-1])
    th[:-1, 1:] = np.maximum(0.5 * t[1:, :-1], th[:-1, 1:])
    th[1:, :-1] = np.maximum(0.5 * t[:-1, 1:], th[1:, :-1])
    th[1:, 1:] = np.maximum(0.5 * t[:-1, :-1], th[1:, 1:])
    th[:, :-1] = np.maximum(0.25 * (th[:, 1:] > 0), th[:, :-1])
    th[:, 1:] = np.maximum(0.25 * (th[:, :-1] > 0), th[:, 1:])
    th[:-1, :] = np.maximum(0.25 * (th[1:, :] > 0), th[:-1, :])
    th[1:, :] = np.maximum(0.25 * (th[:-1, :] > 0), th[1:, :])
    return np.maximum(th[2:-1, 2:-2], th[1:-2, 2:-2]), np.maximum(th[2:-2, 2:-1], th[2:-2, 1:-2])


# ---------------------------------------------------------------------------
# boundary conditions (fvhoe/boundary_conditions.py:60-151,306-402)
# ---------------------------------------------------------------------------


def fv_uniform_meshgen(n, x=(0, 1), y=(0, 1), z=(0, 1), slab_thickness=(0, 0, 0)):
    """Cell centres, and with a slab thickness the ghost-slab centres per side
    (fvhoe/fv.py:12-86): slab "xl" spans slab_thickness[0] cells left of x[0]
    and the interior padded by the other axes' thicknesses."""
    def mesh(m, xx, yy, zz):
        mids = []
        for (a, b), k in zip((xx, yy, zz), m):
            bnd = np.linspace(a, b, k + 1)
            mids.append(0.5 * (bnd[1:] + bnd[:-1]))
        return tuple(np.meshgrid(*mids, indexing="ij"))

    inner = mesh(n, x, y, z)
    if tuple(slab_thickness) == (0, 0, 0):
        return inner
    h = ((x[1] - x[0]) / n[0], (y[1] - y[0]) / n[1], (z[1] - z[0]) / n[2])
    g = tuple(slab_thickness)
    dom = (x, y, z)
    slabs = {}
    for d, dim in enumerate("xyz"):
        for pos in "lr":
            cells, ext = [], []
            for e in range(3):
                lo, hi = dom[e]
                if e == d:
                    cells.append(g[e])
                    ext.append((lo - h[e] * g[e], lo) if pos == "l" else (hi, hi + h[e] * g[e]))
                else:
                    cells.append(n[e] + 2 * g[e])
                    ext.append((lo - h[e] * g[e], hi + h[e] * g[e]))
            slabs[dim + pos] = mesh(cells, *ext)
    return inner, slabs


def trim_slab(slab_coords, gw, axis: int, pos: str):
    """The slab-centre coordinates of one side for ghost widths gw, cut from a
    thicker buffer (fvhoe/boundary_conditions.py:240-304)."""
    buf = (slab_coords["xl"][0].shape[0], slab_coords["yl"][0].shape[1], slab_coords["zl"][0].shape[2])
    X, Y, Z = slab_coords["xyz"[axis] + pos]
    cut = []
    for e in range(3):
        ex = buf[e] - gw[e]
        if ex < 0:
            raise ValueError(f"Cannot apply bc to region of thickness {gw[e]} with a buffer of thickness {buf[e]}")
        if e == axis:
            cut.append(slice(-gw[e], None) if pos == "l" else slice(None, gw[e]))
        else:
            cut.append(slice(ex or None, -ex or None))
    cut = tuple(cut)
    return X[cut], Y[cut], Z[cut]


def apply_bc(u: np.ndarray, gw, bc, values=None, slab_coords=None, t=None) -> np.ndarray:
    """Pad u by gw per axis; slabs filled in x(l,r), y(l,r), z(l,r) order.

    bc: ((xl, xr), (yl, yr), (zl, zr)) with types periodic | outflow |
    reflective | free | dirichlet | ic.  outflow = mirror, reflective =
    mirror with the normal momentum negated, free = nearest interior cell
    (fvhoe/boundary_conditions.py:357-365); dirichlet / ic = the side's value
    in values (same nesting as bc): a per-variable array broadcast over the
    slab, or a callable of the slab centres (x, y, z, plus t for dirichlet)
    (fvhoe/boundary_conditions.py:18-57,334-348)."""
    out = np.pad(u, [(0, 0)] + [(g, g) for g in gw], mode="constant", constant_values=np.nan)
    for i in range(3):
        g = gw[i]
        if g == 0:
            continue
        a = i + 1
        L = out.shape[a]

        def sl(lo, hi, step=None):
            s = [slice(None)] * 4
            s[a] = slice(lo, hi, step)
            return tuple(s)

        for j, pos in enumerate("lr"):
            kind = bc[i][j]
            if kind == "periodic":
                if pos == "l":
                    out[sl(0, g)] = out[sl(L - 2 * g, L - g)]
                    out[sl(L - g, L)] = out[sl(g, 2 * g)]
            elif kind in ("outflow", "reflective"):
                if pos == "l":
                    out[sl(0, g)] = out[sl(g, 2 * g)][sl(None, None, -1)]
                    ghost = sl(0, g)
                else:
                    out[sl(L - g, L)] = out[sl(L - 2 * g, L - g)][sl(None, None, -1)]
                    ghost = sl(L - g, L)
                if kind == "reflective":
                    gs = list(ghost)
                    gs[0] = a
                    out[tuple(gs)] *= -1.0
            elif kind == "free":
                if pos == "l":
                    out[sl(0, g)] = out[sl(g, g + 1)]
                else:
                    out[sl(L - g, L)] = out[sl(L - g - 1, L - g)]
            elif kind in ("dirichlet", "ic"):
                view = out[sl(0, g)] if pos == "l" else out[sl(L - g, L)]
                X, Y, Z = trim_slab(slab_coords, gw, i, pos)
                val = values[i][j]
                if callable(val):
                    kw = dict(x=X, y=Y, z=Z)
                    if kind == "dirichlet":
                        kw["t"] = t
                    view[...] = val(**kw)
                elif isinstance(val, np.ndarray):
                    view[...] = np.ones_like(view) * val.reshape(-1, 1, 1, 1)
                else:
                    raise ValueError(f"Invalid boundary of type '{val}'.")
            elif kind is None:
                pass
            else:
                raise ValueError(f"Unknown bc type '{kind}'")
    return out


# ---------------------------------------------------------------------------
# one stage (fvhoe/solver.py:422-792)
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class Scheme:
    """Everything one stage needs; mirrors EulerSolver's ctor arguments
    (fvhoe/solver.py:42-84) that reach the hot path."""

    n: Tuple[int, int, int]
    p: Tuple[int, int, int]
    domain: Tuple[Tuple[float, float], ...] = ((0, 1), (0, 1), (0, 1))
    gamma: float = 5 / 3
    CFL: float = 0.8
    fixed_dt: Optional[float] = None
    bc: tuple = (("periodic", "periodic"),) * 3
    riemann_solver: str = "hllc"
    a_posteriori: bool = False
    slope_limiter: Optional[str] = "minmod"
    force_trouble: bool = False
    NAD: float = 1e-2
    NAD_mode: str = "global"
    NAD_range: str = "relative"
    NAD_vars: Optional[tuple] = None
    PAD: Optional[dict] = None
    SED: bool = True
    SED_tolerance: float = 1e-10
    convex: bool = False
    csq_floor: float = 1e-10
    bc_values: tuple = None  # per side, for dirichlet / ic sides (same nesting as bc)
    slab_buffer_size: int = 30
    t: float = 0.0  # the step's start time: every stage fills ghosts at it (fvhoe/solver.py:784)
    bulk: str = "fp64"  # "fp32": float32 intermediates (the FP32 bulk form of the composed stage)

    def __post_init__(self):
        self.h = tuple((b - a) / n for (a, b), n in zip(self.domain, self.n))
        self.active = tuple(not (n == 1 and p == 0) for n, p in zip(self.n, self.p))
        self.dims = "".join(d for d, f in zip("xyz", self.active) if f)
        self.ndim = len(self.dims)
        if self.PAD is None:
            self.PAD = {"rho": (0.0, np.inf), "P": (0.0, np.inf)}  # fvhoe/solver.py:392-396
        self.bc = tuple(tuple(b) if isinstance(b, (tuple, list)) else (b, b) for b in self.bc)
        self.slab_coords = None
        if any(k in ("dirichlet", "ic") for b in self.bc for k in b):
            sizes = tuple(self.slab_buffer_size if a else 0 for a in self.active)
            _, self.slab_coords = fv_uniform_meshgen(self.n, *self.domain, slab_thickness=sizes)


def interpolate_primitives_from_conservatives(s: Scheme, u, p, gw=(0, 0, 0), fv_average=True):
    """fvhoe/solver.py:753-792."""
    cost = [(2 if fv_average else 1) * stencil_reach(pi) for pi in p]
    g = tuple(gw[i] + cost[i] for i in range(3))
    u_bc = apply_bc(u, g, s.bc, s.bc_values, s.slab_coords, s.t)
    if s.bulk == "fp32":
        u_bc = u_bc.astype(np.float32)
    w = compute_primitives(interpolate_cell_centers(u_bc, p), s.gamma)
    return interpolate_fv_averages(w, p) if fv_average else w


def hydrofluxes(s: Scheme, u, p, slope_limiter=None):
    """dt, (F, G, H) of degree-p reconstruction (fvhoe/solver.py:456-651)."""
    if slope_limiter is None:
        limit = False
    elif max(p) <= 1:
        limit = True
    else:
        raise ValueError("Slope limiting is only implemented for degree 1 polynomial interpolation.")
    pmax = max(p)
    cm = stencil_reach(pmax)
    gw = max(cm + 1, 2 * cm)
    gws = tuple(gw if a else 0 for a in s.active)
    w_bc = interpolate_primitives_from_conservatives(s, u, p, gws)
    dt = s.fixed_dt if s.fixed_dt is not None else hydro_dt(
        w_bc, min(s.h), s.ndim, s.CFL, s.gamma, s.csq_floor)
    rs = RIEMANN[s.riemann_solver]
    # transverse (c) sweep order per face direction (fvhoe/solver.py:522-568)
    order = {0: (3, 2), 1: (1, 3), 2: (2, 1)}
    tr_order = {0: (2, 3), 1: (1, 3), 2: (1, 2)}
    out = []
    for d in range(3):
        if not s.active[d]:
            out.append(None)
            continue
        a = d + 1
        wt = w_bc
        for ax in order[d]:
            wt = conservative_interpolation(wt, p[ax - 1], ax, "c")
        if limit and p[d] == 1:
            wl_face, wr_face = MUSCL_interpolations(wt, a, slope_limiter)
        else:
            wl_face = conservative_interpolation(wt, p[d], a, "l")
            wr_face = conservative_interpolation(wt, p[d], a, "r")
        gamma, csq_floor = s.gamma, s.csq_floor
        if wt.dtype == np.float32:  # the float32 solve takes gamma and the floor rounded to float32
            gamma, csq_floor = np.float32(gamma), np.float32(csq_floor)
        f = rs(_shift(wr_face, a, 0, 1), _shift(wl_face, a, 1, 0), gamma, "xyz"[d], csq_floor)
        for ax in tr_order[d]:
            f = transverse_reconstruction(f, p[ax - 1], ax)
        excess = [0, 0, 0, 0]
        for e in range(3):
            if s.active[e]:
                excess[e + 1] = gw - stencil_reach(p[e]) - 1 if e == d else gw - 2 * stencil_reach(p[e])
        f = f[tuple(slice(x or None, -x or None) for x in excess)]
        out.append(f)
    return dt, tuple(out)


def euler_equation(s: Scheme, F, G, H):
    """dudt = 0.0 + -(1/hx)(F[1:]-F[:-1]) + ... in x, y, z order (fvhoe/solver.py:734-751)."""
    dudt = 0.0
    F, G, H = (None if f is None else f.astype(np.float64, copy=False) for f in (F, G, H))
    if s.active[0]:
        dudt = dudt + -(1 / s.h[0]) * (F[:, 1:] - F[:, :-1])
    if s.active[1]:
        dudt = dudt + -(1 / s.h[1]) * (G[:, :, 1:] - G[:, :, :-1])
    if s.active[2]:
        dudt = dudt + -(1 / s.h[2]) * (H[:, :, :, 1:] - H[:, :, :, :-1])
    return dudt


@dataclasses.dataclass
class StageLog:
    trouble: Optional[np.ndarray] = None
    NAD_mag: Optional[np.ndarray] = None
    PAD_mag: Optional[np.ndarray] = None


def revise_fluxes(s: Scheme, u, F, G, H, dt, log: Optional[StageLog] = None):
    """A-posteriori check and MUSCL fallback blend (fvhoe/solver.py:653-732).
    Returns revised (F, G, H)."""
    ustar = u + dt * euler_equation(s, F, G, H)
    gw = tuple(3 if a else 0 for a in s.active)
    w = interpolate_primitives_from_conservatives(s, u, s.p, gw)
    ws = interpolate_primitives_from_conservatives(s, ustar, s.p, gw)
    trouble, nad_mag, pad_mag = detect_troubled_cells(
        w[:5], ws[:5], s.dims, s.NAD, s.NAD_mode, s.NAD_range, s.NAD_vars, s.PAD, s.SED, s.SED_tolerance)
    if s.force_trouble:
        trouble = np.ones_like(trouble, dtype=bool)
    if log is not None:
        log.trouble, log.NAD_mag, log.PAD_mag = trouble, nad_mag, pad_mag
    if np.any(trouble):
        p1 = tuple(int(pi > 0) for pi in s.p)
        _, (Fl, Gl, Hl) = hydrofluxes(s, u, p1, slope_limiter=s.slope_limiter)
        periodic = tuple(b == ("periodic", "periodic") for b in s.bc)
        tx, ty, tz = broadcast_to_troubled_interfaces(trouble, s.dims, s.convex, periodic)
        if F is not None:
            F = (1 - tx) * F + tx * Fl
        if G is not None:
            G = (1 - ty) * G + ty * Gl
        if H is not None:
            H = (1 - tz) * H + tz * Hl
    return F, G, H


def hydrodynamics(s: Scheme, u, log: Optional[StageLog] = None):
    """Per-stage RHS: (dt_stage, dudt) (fvhoe/solver.py:429-454)."""
    dt, (F, G, H) = hydrofluxes(s, u, s.p)
    if s.a_posteriori:
        F, G, H = revise_fluxes(s, u, F, G, H, dt, log)
    return dt, euler_equation(s, F, G, H)


# ---------------------------------------------------------------------------
# time integration (fvhoe/ode.py:11-22,183-193,245-298)
# ---------------------------------------------------------------------------


def hit_target(t: float, dt: float, target: Optional[float]) -> float:
    """fvhoe/ode.py:11-22."""
    return target - t if (target is not None and target - t < dt) else dt


def step(s: Scheme, u, t: float, integrator: str = "ssprk3", target: Optional[float] = None,
         logs: Optional[list] = None):
    """One step of the reference's k-form steppers; returns (dt, unext)."""
    s.t = t

    def f(uu):
        lg = StageLog() if logs is not None else None
        out = hydrodynamics(s, uu, lg)
        if logs is not None:
            logs.append(lg)
        return out

    if integrator == "euler":
        dt, k0 = f(u)
        dt = hit_target(t, dt, target)
        return dt, u + dt * k0
    if integrator == "ssprk2":
        dt, k0 = f(u)
        dt = hit_target(t, dt, target)
        u1 = u + dt * k0
        _, k1 = f(u1)
        return dt, 0.5 * u + 0.5 * (u1 + dt * k1)
    if integrator == "ssprk3":
        dt, k0 = f(u)
        dt = hit_target(t, dt, target)
        _, k1 = f(u + dt * k0)
        _, k2 = f(u + 0.25 * dt * k0 + 0.25 * dt * k1)
        return dt, u + (1 / 6) * dt * (k0 + k1 + 4 * k2)
    if integrator == "rk4":
        dt, k0 = f(u)
        dt = hit_target(t, dt, target)
        _, k1 = f(u + 0.5 * dt * k0)
        _, k2 = f(u + 0.5 * dt * k1)
        _, k3 = f(u + dt * k2)
        return dt, u + (1 / 6) * dt * (k0 + 2 * k1 + 2 * k2 + k3)
    raise ValueError(integrator)


def integrate(s: Scheme, u0, T: Optional[float] = None, n: Optional[int] = None,
              integrator: str = "ssprk3", logs: Optional[list] = None):
    """Run to time T (exact, fvhoe/ode.py:94-174) or for n steps.
    Returns (t, u, step_count)."""
    u = u0.copy()
    t = 0.0
    steps = 0
    if T is None:
        for _ in range(n):
            dt, u = step(s, u, t, integrator, None, logs)
            t += dt
            steps += 1
        return t, u, steps
    while t < T:
        dt, u = step(s, u, t, integrator, T, logs)
        t += dt
        steps += 1
    return t, u, steps


def primitive_snapshot(s: Scheme, u):
    """snapshot()['w']: primitive fv averages of the interior (fvhoe/solver.py:845-849)."""
    return interpolate_primitives_from_conservatives(s, u, s.p, (0, 0, 0))


def l1err(x, y) -> float:
    """tests/utils.py:4-12."""
    return float(np.mean(np.abs(x - y)))
