"""Dirichlet / "ic" boundaries in the oracle and in the host side of
BoundaryCondition (fvhoe/boundary_conditions.py:18-57,240-348, fvhoe/fv.py:12-86),
pinned by the reference's own cases (tests/test_boundary_conditions.py:92-230)
and by closed forms of the slab geometry."""

import numpy as np
import pytest

from fvhoe_amd.boundary_conditions import BoundaryCondition
from fvhoe_amd.fv import fv_uniform_meshgen
from fvhoe_amd.initial_conditions import ShuOsher1D, primitives_to_conservatives
from oracle import fvhoe_oracle as O


def _f(x, y, z, t):
    d = np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z) * np.sin(2 * np.pi * t)
    return np.stack([d, 3 * d, 4 * d, 5 * d, 2 * d])


def test_slab_centres_closed_form():
    n, g = (6, 5, 4), (2, 3, 1)
    dom = ((0.0, 1.5), (-1.0, 1.0), (0.0, 2.0))
    h = [(b - a) / k for (a, b), k in zip(dom, n)]
    (X, Y, Z), sl = O.fv_uniform_meshgen(n, *dom, slab_thickness=g)
    for d, dim in enumerate("xyz"):
        for pos in "lr":
            C = sl[dim + pos]
            shape = tuple(g[e] if e == d else n[e] + 2 * g[e] for e in range(3))
            assert C[0].shape == shape
            for e in range(3):
                k = np.arange(shape[e])
                lo, hi = dom[e]
                if e == d:
                    want = lo - (g[e] - k - 0.5) * h[e] if pos == "l" else hi + (k + 0.5) * h[e]
                else:
                    want = lo + (k - g[e] + 0.5) * h[e]
                got = np.moveaxis(C[e], e, 0).reshape(shape[e], -1)
                assert np.allclose(got, want[:, None], rtol=0, atol=1e-14)
    # the package's copy is the same arithmetic
    (X2, Y2, Z2), sl2 = fv_uniform_meshgen(n, *dom, slab_thickness=g)
    assert np.array_equal(X, X2) and np.array_equal(Z, Z2)
    for k in sl:
        for a, b in zip(sl[k], sl2[k]):
            assert np.array_equal(a, b)


def test_dirichlet_f_of_xyzt(N: int = 32, gw: int = 8, t: float = 0.5):
    """The reference's test_dirichlet_f_of_xyzt: a callable on every side gives f on the padded mesh
    (h = 1/32 keeps every centre exact in binary, so the two meshes agree bit for bit)."""
    h = 1 / N
    Xa, Ya, Za = O.fv_uniform_meshgen((N + 2 * gw,) * 3, *((-h * gw, 1 + h * gw),) * 3)
    (X, Y, Z), slabs = O.fv_uniform_meshgen((N, N, N), slab_thickness=(gw, gw, gw))
    kinds = (("dirichlet", "dirichlet"),) * 3
    vals = ((_f, _f),) * 3
    out = O.apply_bc(_f(X, Y, Z, t), (gw, gw, gw), kinds, vals, slabs, t)
    assert np.array_equal(out, _f(Xa, Ya, Za, t))


def test_dirichlet_corners_follow_the_last_axis():
    """x Dirichlet then y outflow: the y mirror carries the x slab into the corners;
    x outflow then y Dirichlet: the y slab decides the corners."""
    n, g = (8, 6, 1), (3, 2, 0)
    rng = np.random.default_rng(0)
    u = rng.random((5, *n))
    (_, _, _), slabs = O.fv_uniform_meshgen(n, slab_thickness=(4, 4, 0))
    c = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    out = O.apply_bc(u, g, (("dirichlet", "outflow"), ("outflow", "outflow"), ("periodic", "periodic")),
                     ((c, None), (None, None), (None, None)), slabs)
    assert np.array_equal(out[:, :3, :, 0], np.broadcast_to(c[:, None, None], (5, 3, 10)))
    assert np.array_equal(out[:, 3:-3, :2], out[:, 3:-3, 2:4][:, :, ::-1])  # plain mirror in y
    fy = lambda x, y, z, t: np.stack([x + 10 * y + k for k in range(5)])  # noqa: E731
    out = O.apply_bc(u, g, (("outflow", "outflow"), ("dirichlet", "reflective"), ("periodic", "periodic")),
                     ((None, None), (fy, None), (None, None)), slabs, 0.0)
    X, Y, Z = O.trim_slab(slabs, g, 1, "l")
    assert np.array_equal(out[:, :, :2], fy(X, Y, Z, 0.0))  # corners included: the y slab spans x ghosts
    mirror = out[:, :, -4:-2][:, :, ::-1].copy()
    mirror[2] *= -1
    assert np.array_equal(out[:, :, -2:], mirror)


def test_slab_buffer_too_thin_raises():
    (_, _, _), slabs = O.fv_uniform_meshgen((8, 1, 1), slab_thickness=(2, 0, 0))
    with pytest.raises(ValueError, match="buffer of thickness 2"):
        O.apply_bc(np.ones((5, 8, 1, 1)), (3, 0, 0), (("ic", "ic"), (None, None), (None, None)),
                   ((np.ones(5), np.ones(5)), (None, None), (None, None)), slabs)


@pytest.mark.parametrize("p", [0, 1, 3])
def test_ic_equals_constant_dirichlet_shu_osher(p, N=100, gamma=1.4, T=1.8):
    """The reference's test_ic_bc along x: "ic" ghosts (pointwise IC at the slab centres) and constant
    Dirichlet ghosts give the same pressure at T = 1.8 (the right-hand difference is a density
    perturbation at rest, which never leaves the ghosts)."""
    w0 = ShuOsher1D(dim="x").build_ic()
    point = lambda x, y, z: primitives_to_conservatives(w0(x, y, z), gamma)  # noqa: E731
    s_kw = dict(n=(N, 1, 1), p=(p, 0, 0), domain=((0, 10), (0, 1), (0, 1)), gamma=gamma, riemann_solver="hllc",
                a_posteriori=p > 0, slope_limiter="minmod")
    ends = (np.array([3.857143, 3.857143 * 2.629369, 0, 0, 10.33333 / (gamma - 1) + 0.5 * 3.857143 * 2.629369 ** 2]),
            np.array([1 + 0.2 * np.sin(5 * 5), 0, 0, 0, 1 / (gamma - 1)]))
    s_dir = O.Scheme(bc=(("dirichlet", "dirichlet"), (None, None), (None, None)),
                     bc_values=(ends, (None, None), (None, None)), **s_kw)
    s_ic = O.Scheme(bc=(("ic", "ic"), (None, None), (None, None)),
                    bc_values=((point, point), (None, None), (None, None)), **s_kw)
    X, Y, Z = O.fv_uniform_meshgen((N, 1, 1), (0, 10))
    pw, ww = np.polynomial.legendre.leggauss(-(-(p + 1) // 2))
    h = 10 / N
    u0 = sum(w / 2 * point(X + x / 2 * h, Y, Z) for x, w in zip(pw, ww))
    integ = {0: "euler", 1: "ssprk2"}.get(p, "ssprk3")
    _, ud, _ = O.integrate(s_dir, u0, T, integrator=integ)
    _, ui, _ = O.integrate(s_ic, u0, T, integrator=integ)
    Pd = O.primitive_snapshot(s_dir, ud)[4]
    Pi = O.primitive_snapshot(s_ic, ui)[4]
    assert O.l1err(Pd, Pi) < 1e-14
    # and the two boundary kinds did differ in the ghosts (density on the right)
    assert not np.array_equal(O.apply_bc(u0, (2, 0, 0), s_dir.bc, s_dir.bc_values, s_dir.slab_coords),
                              O.apply_bc(u0, (2, 0, 0), s_ic.bc, s_ic.bc_values, s_ic.slab_coords))


def test_host_slab_values_match_the_oracle():
    n = (10, 7, 1)
    (_, _, _), slabs = fv_uniform_meshgen(n, slab_thickness=(6, 6, 0))
    c = np.arange(5.0) + 0.5
    bc = BoundaryCondition(x=("dirichlet", "ic"), y=("outflow", "dirichlet"), z=None,
                           x_value=(c, lambda x, y, z: np.stack([x * y + k for k in range(5)])),
                           y_value=(None, _f), slab_coords=slabs)
    gw = (4, 3, 0)
    for axis, pos, val, kind in ((0, "l", c, "dirichlet"), (0, "r", bc.x_value[1], "ic"), (1, "r", _f, "dirichlet")):
        X, Y, Z = O.trim_slab(slabs, gw, axis, pos)
        got = bc.slab_values(gw, axis, pos, 5, t=0.3)
        if callable(val):
            want = val(x=X, y=Y, z=Z, t=0.3) if kind == "dirichlet" else val(x=X, y=Y, z=Z)
        else:
            want = np.ones((5, *X.shape)) * val.reshape(-1, 1, 1, 1)
        assert np.array_equal(got, want)
    assert bc.time_dependent() and bc.has_slabs()
    assert not BoundaryCondition(x="ic", x_value=c, slab_coords=slabs).time_dependent()


def test_host_boundary_condition_errors():
    with pytest.raises(ValueError, match="list"):
        BoundaryCondition(x=["periodic", "periodic"])
    with pytest.raises(ValueError, match="Unknown bc type"):
        BoundaryCondition(x="mirror")
    (_, _, _), slabs = fv_uniform_meshgen((8, 1, 1), slab_thickness=(2, 0, 0))
    bc = BoundaryCondition(x="dirichlet", x_value="3", slab_coords=slabs)
    with pytest.raises(ValueError, match="Invalid boundary of type"):
        bc.slab_values((2, 0, 0), 0, "l", 5)
    with pytest.raises(ValueError, match="buffer of thickness 2"):
        bc.slab_values((3, 0, 0), 0, "l", 5)
    with pytest.raises(ValueError, match="slab_coords"):
        BoundaryCondition(x="ic", x_value=np.ones(5)).slab_values((2, 0, 0), 0, "l", 5)
