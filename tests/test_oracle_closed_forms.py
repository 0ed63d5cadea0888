"""The oracle against what the mathematics fixes, independently of the golden
vectors: polynomial exactness of the stencils in exact rationals, brute-force
sweeps, limiters and ghost fills on tiny arrays, closed forms of the
thermodynamics and the CFL step, and the textbook properties of the Riemann
solvers (consistency, supersonic upwinding, mirror symmetry).  Each check is
chosen so that a plausible slip (a wrong weight, a transposed sweep axis, a
dropped term, a sign) fails it."""

from fractions import Fraction

import numpy as np
import pytest

from oracle import fvhoe_oracle as O

GAMMA = 1.4


# --------------------------------------------------------------- stencils
def _cell_average_of_monomial(q: int, a: Fraction, b: Fraction) -> Fraction:
    return (b ** (q + 1) - a ** (q + 1)) / ((q + 1) * (b - a))


@pytest.mark.parametrize("p", range(0, 9))
def test_stencils_are_exact_for_polynomials_of_degree_p(p):
    """Cells [k-1/2, k+1/2] of width 1 around the stencil centre 0: the l / c / r weights map the
    averages of any polynomial of degree <= p (<= p-1 for the centre of odd p, whose outer taps are zero)
    to its point values at -1/2, 0, +1/2; the transverse weights map point values at the cell centres to
    the average over the centre cell."""
    r = O.stencil_reach(p)
    cells = range(-r, r + 1)
    for pos, x0 in (("l", Fraction(-1, 2)), ("c", Fraction(0)), ("r", Fraction(1, 2))):
        w = O.interpolation_weights_exact(p, pos)
        assert len(w) == 2 * r + 1
        qmax = p - 1 if (pos == "c" and p % 2) else p
        for q in range(qmax + 1):
            got = sum(wi * _cell_average_of_monomial(q, k - Fraction(1, 2), k + Fraction(1, 2))
                      for wi, k in zip(w, cells))
            assert got == x0 ** q, (p, pos, q)
        # and the float64 table is each rational correctly rounded
        assert list(O.interpolation_weights(p, pos)) == [float(f) for f in w]
    w = O.transverse_weights_exact(p)
    qmax = p - 1 if p % 2 else p
    for q in range(qmax + 1):
        got = sum(wi * Fraction(k) ** q for wi, k in zip(w, cells))
        assert got == _cell_average_of_monomial(q, Fraction(-1, 2), Fraction(1, 2)), (p, q)


def test_left_and_right_stencils_mirror():
    for p in range(9):
        assert O.interpolation_weights_exact(p, "r") == tuple(reversed(O.interpolation_weights_exact(p, "l")))


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_stencil_sweep_brute_force(axis):
    rng = np.random.default_rng(axis)
    u = rng.standard_normal((2, 7, 6, 8))
    w = O.interpolation_weights(3, "l")
    got = O.stencil_sweep(u, w, axis)
    k = len(w)
    want = np.zeros(got.shape)
    for idx in np.ndindex(got.shape):
        acc = 0.0
        for t in range(k):
            src = list(idx)
            src[axis] += t
            acc = acc + w[t] * u[tuple(src)]
        want[idx] = acc
    assert np.array_equal(got, want)


def test_cell_centres_and_averages_of_a_polynomial():
    """x-y-z c-sweeps of exact cell averages of a degree-2 polynomial give its centre values; the
    transverse sweeps of centre values give the averages back (p = 2: exact for degree 2)."""
    n = 8
    xs = np.arange(n, dtype=np.float64)
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")

    def avg(x, y, z):  # average over the unit cell of f = 1 + x^2 - 2xy + 3z^2 + y
        return 1 + (x ** 2 + 1 / 12) - 2 * x * y + 3 * (z ** 2 + 1 / 12) + y

    def point(x, y, z):
        return 1 + x ** 2 - 2 * x * y + 3 * z ** 2 + y

    u = avg(X, Y, Z)[None]
    c = O.interpolate_cell_centers(u, (2, 2, 2))
    assert np.allclose(c[0], point(X, Y, Z)[1:-1, 1:-1, 1:-1], rtol=0, atol=1e-11)
    a = O.interpolate_fv_averages(point(X, Y, Z)[None], (2, 2, 2))
    assert np.allclose(a[0], avg(X, Y, Z)[1:-1, 1:-1, 1:-1], rtol=0, atol=1e-11)


# --------------------------------------------------------- thermodynamics
def _states(shape, seed):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((5, *shape))
    w[0] = 0.2 + rng.random(shape)
    w[4] = 0.2 + rng.random(shape)
    return w


def test_primitives_conservatives_round_trip_and_closed_form():
    w = _states((5, 4, 3), 0)
    u = O.compute_conservatives(w, GAMMA)
    assert np.allclose(u[1:4], w[0] * w[1:4], rtol=1e-15, atol=0)
    assert np.allclose(u[4], w[4] / (GAMMA - 1) + 0.5 * w[0] * (w[1] ** 2 + w[2] ** 2 + w[3] ** 2), rtol=1e-14)
    assert np.allclose(O.compute_primitives(u, GAMMA), w, rtol=1e-12, atol=1e-14)


def test_sound_speed_and_cfl_step_closed_forms():
    w = _states((6, 5, 1), 1)
    c = O.compute_sound_speed(w, GAMMA, 1e-10)
    assert np.allclose(c, np.sqrt(GAMMA * w[4] / w[0]), rtol=1e-15)
    w2 = w.copy()
    w2[4, 0, 0, 0] = 0.0  # floored
    assert O.compute_sound_speed(w2, GAMMA, 1e-10)[0, 0, 0] == np.sqrt(1e-10)
    dt = O.hydro_dt(w, 0.01, 2, 0.8, GAMMA, 1e-10)
    speed = np.abs(w[1]) + np.abs(w[2]) + np.abs(w[3]) + 2 * c
    assert np.isclose(dt, 0.8 * 0.01 / speed.max(), rtol=1e-15)


# ------------------------------------------------------------ Riemann solvers
def _physical_flux(w, d):
    u = O.compute_conservatives(w, GAMMA)
    f = w[d] * u
    f[d] = f[d] + w[4]
    f[4] = f[4] + w[d] * w[4]
    return f


@pytest.mark.parametrize("rs", ["llf", "hllc"])
@pytest.mark.parametrize("dim", ["x", "y", "z"])
def test_riemann_consistency(rs, dim):
    w = _states((16, 1, 1), 2)
    got = O.RIEMANN[rs](w, w, GAMMA, dim, 1e-10)
    assert np.allclose(got, _physical_flux(w, O.AXES[dim]), rtol=1e-12, atol=1e-12)


def test_llf_textbook_form():
    """Toro (p.331): F = (Fl + Fr)/2 - max(|vl| + cl, |vr| + cr)/2 (ur - ul)."""
    wl, wr = _states((16, 1, 1), 3), _states((16, 1, 1), 4)
    d = 2
    cl = np.sqrt(GAMMA * wl[4] / wl[0])
    cr = np.sqrt(GAMMA * wr[4] / wr[0])
    s = np.maximum(np.abs(wl[d]) + cl, np.abs(wr[d]) + cr)
    want = 0.5 * (_physical_flux(wl, d) + _physical_flux(wr, d)) - 0.5 * s * (
        O.compute_conservatives(wr, GAMMA) - O.compute_conservatives(wl, GAMMA))
    assert np.allclose(O.llf(wl, wr, GAMMA, "y", 1e-10), want, rtol=1e-12, atol=1e-12)


def test_hllc_supersonic_states_upwind():
    wl, wr = _states((16, 1, 1), 5), _states((16, 1, 1), 6)
    fast = wl.copy(), wr.copy()
    for w in fast:
        w[1] = 10.0  # Sl > 0: the left state's flux
    assert np.allclose(O.hllc(*fast, GAMMA, "x", 1e-10), _physical_flux(fast[0], 1), rtol=1e-12, atol=1e-12)
    slow = wl.copy(), wr.copy()
    for w in slow:
        w[1] = -10.0  # Sr < 0: the right state's flux
    assert np.allclose(O.hllc(*slow, GAMMA, "x", 1e-10), _physical_flux(slow[1], 1), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("rs", ["llf", "hllc"])
def test_riemann_mirror_symmetry(rs):
    """Reflecting the problem (swap sides, negate the normal velocity) negates the mass, transverse
    momentum and energy fluxes and keeps the normal momentum flux."""
    wl, wr = _states((16, 1, 1), 7), _states((16, 1, 1), 8)
    ml, mr = wr.copy(), wl.copy()
    ml[1] *= -1
    mr[1] *= -1
    f = O.RIEMANN[rs](wl, wr, GAMMA, "x", 1e-10)
    g = O.RIEMANN[rs](ml, mr, GAMMA, "x", 1e-10)
    assert np.allclose(g[[0, 2, 3, 4]], -f[[0, 2, 3, 4]], rtol=1e-12, atol=1e-13)
    assert np.allclose(g[1], f[1], rtol=1e-12, atol=1e-13)


# --------------------------------------------------------------- limiters
def test_limiters_brute_force():
    rng = np.random.default_rng(9)
    dl = rng.standard_normal(200)
    dr = rng.standard_normal(200)
    dl[:5] = 0.0
    for a, b, mm, mc in zip(dl, dr, O.minmod(dl, dr), O.moncen(dl, dr)):
        den = max(a, 1e-16) if a > 0 else min(a, -1e-16)
        want_mm = min(max(b / den, 0.0), 1.0) * a
        assert mm == want_mm
        if a * b >= 0:
            dc = 0.5 * (a + b)
            want_mc = np.sign(dc) * min(abs(2 * a), abs(2 * b), abs(dc))
        else:
            want_mc = 0.0
        assert mc == want_mc


def test_muscl_faces_brute_force():
    rng = np.random.default_rng(10)
    u = rng.standard_normal((1, 9, 1, 1))
    left, right = O.MUSCL_interpolations(u, 1, "minmod")
    for i in range(1, 8):
        a, b, c = u[0, i - 1, 0, 0], u[0, i, 0, 0], u[0, i + 1, 0, 0]
        s = O.minmod(np.array(b - a), np.array(c - b))
        assert left[0, i - 1, 0, 0] == b - 0.5 * s and right[0, i - 1, 0, 0] == b + 0.5 * s


# -------------------------------------------------------------- ghost fills
@pytest.mark.parametrize("kind", ["periodic", "outflow", "reflective", "free"])
def test_ghost_fill_index_maps(kind):
    n, g = 6, 3
    rng = np.random.default_rng(11)
    u = rng.standard_normal((5, n, 1, 1))
    out = O.apply_bc(u, (g, 0, 0), ((kind, kind), ("periodic", "periodic"), ("periodic", "periodic")))
    for j in range(n + 2 * g):
        c = j - g
        if 0 <= c < n:
            src, flip = c, False
        elif kind == "periodic":
            src, flip = c % n, False
        elif kind == "free":
            src, flip = min(max(c, 0), n - 1), False
        else:
            src, flip = (-1 - c if c < 0 else 2 * n - 1 - c), kind == "reflective"
        want = u[:, src, 0, 0].copy()
        if flip:
            want[1] = -want[1]
        assert np.array_equal(out[:, j, 0, 0], want), (kind, j)


# ------------------------------------------------- central difference, fluxes
@pytest.mark.parametrize("axis", [0, 1, 2, 3])
def test_central_difference_of_a_linear_field_is_its_slope(axis):
    """0.5 (u[i+1] - u[i-1]) / h of u = a + b x_i (x_i = i h, all values dyadic) is b exactly; without h it
    is b h.  A wrong shift, sign or scale fails."""
    h, a, b = 0.125, 3.0, -1.5
    shape = [4, 5, 6, 7]
    x = np.arange(shape[axis]) * h
    u = np.broadcast_to((a + b * x).reshape([-1 if k == axis else 1 for k in range(4)]), shape).copy()
    d = O.second_order_central_difference(u, axis, h)
    want = list(shape)
    want[axis] -= 2
    assert d.shape == tuple(want)
    assert np.all(d == b)
    assert np.all(O.second_order_central_difference(u, axis) == b * h)


@pytest.mark.parametrize("dim", ["x", "y", "z"])
def test_physical_flux_closed_forms(dim):
    """Euler fluxes (Toro 2009, eq. 3.3): at rest only the pressure pushes the normal momentum; in motion
    F = (rho v, rho v vx + P dx, rho v vy + P dy, rho v vz + P dz, v (E + P)) up to rounding; without
    pressure every variable is advected."""
    d = O.AXES[dim]
    rng = np.random.default_rng(31)
    w = np.empty((6, 40))
    w[0] = 0.5 + rng.random(40)
    w[1:4] = 0.0
    w[4] = 0.2 + rng.random(40)
    w[5] = rng.random(40)
    u = O.compute_conservatives(w, GAMMA)
    u[5] = w[0] * w[5]
    f = O.compute_fluxes(w, u, dim)
    for q in range(6):
        assert np.all(f[q] == (w[4] if q == d else 0.0)), q
    w[1:4] = rng.standard_normal((3, 40))
    u = O.compute_conservatives(w, GAMMA)
    u[5] = w[0] * w[5]
    f = O.compute_fluxes(w, u, dim)
    v, rho, P, E = w[d], w[0], w[4], u[4]
    want = [rho * v] + [rho * v * w[k] + (P if k == d else 0.0) for k in (1, 2, 3)] + [v * (E + P), v * u[5]]
    for q in range(6):
        assert np.allclose(f[q], want[q], rtol=1e-14, atol=1e-14), q
    g = O.compute_fluxes(w, u, dim, include_pressure=False)
    assert np.array_equal(g[0], v * rho) and all(np.array_equal(g[q], v * u[q]) for q in range(1, 6))
