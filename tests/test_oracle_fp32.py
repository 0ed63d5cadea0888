"""The oracle's FP32 bulk form of the composed stage (float32 intermediates; no
reference counterpart).  Pinned by scalar float32 brute force for the sweeps,
by the float64 recovery for the primitives (one rounding to float32), by the
float64 Riemann solves (float32 rounding only), and by whole float32 stages
against the float64 stage, which the golden vectors pin bit for bit."""

import numpy as np
import pytest

from oracle import fvhoe_oracle as O


def _f32(x):
    return np.float32(x)


@pytest.mark.parametrize("p", [1, 2, 3, 7])
@pytest.mark.parametrize("pos", ["l", "c", "r", "t"])
def test_float32_sweep_is_scalar_float32_arithmetic(p, pos):
    rng = np.random.default_rng(p)
    u = rng.standard_normal((2, 12, 3, 2)).astype(np.float32)
    w64 = O.transverse_weights(p) if pos == "t" else O.interpolation_weights(p, pos)
    got = O.transverse_reconstruction(u, p, 1) if pos == "t" else O.conservative_interpolation(u, p, 1, pos)
    assert got.dtype == np.float32
    want = np.empty_like(got)
    for v, i, j, k in np.ndindex(got.shape):
        acc = _f32(0.0)
        for t, w in enumerate(w64):
            acc = _f32(acc + _f32(_f32(w) * u[v, i + t, j, k]))
        want[v, i, j, k] = acc
    assert np.array_equal(got.view(np.int32), want.view(np.int32))
    # within float32 rounding of the float64 sweep
    ref = (O.transverse_reconstruction(u.astype(np.float64), p, 1) if pos == "t"
           else O.conservative_interpolation(u.astype(np.float64), p, 1, pos))
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-6)


def test_float32_primitives_are_the_float64_recovery_rounded():
    rng = np.random.default_rng(1)
    w = np.stack([0.5 + rng.random((4, 3, 2)), rng.standard_normal((4, 3, 2)), rng.standard_normal((4, 3, 2)),
                  rng.standard_normal((4, 3, 2)), 0.5 + rng.random((4, 3, 2))])
    u32 = O.compute_conservatives(w, 1.4).astype(np.float32)
    got = O.compute_primitives(u32, 1.4)
    assert got.dtype == np.float32
    # closed form on the float32 inputs, evaluated in float64 and rounded once
    u = u32.astype(np.float64)
    v = u[1:4] / u[0]
    P = 0.4 * (u[4] - 0.5 * ((u[1] * v[0] + u[2] * v[1]) + u[3] * v[2]))
    assert np.array_equal(got[1:4], v.astype(np.float32))
    assert np.array_equal(got[4], P.astype(np.float32))


@pytest.mark.parametrize("rs", ["llf", "hllc"])
def test_float32_riemann_solves_round_only(rs):
    rng = np.random.default_rng(2)
    shape = (5, 64, 1, 1)

    def state():
        w = rng.standard_normal(shape)
        w[0] = 0.3 + rng.random(shape[1:])
        w[4] = 0.3 + rng.random(shape[1:])
        return w

    wl, wr = state(), state()
    f64 = O.RIEMANN[rs](wl, wr, 1.4, "x", 1e-10)
    f32 = O.RIEMANN[rs](wl.astype(np.float32), wr.astype(np.float32), np.float32(1.4), "x", np.float32(1e-10))
    assert f32.dtype == np.float32
    scale = np.abs(f64).max()
    assert np.abs(f32 - f64).max() < 1e-5 * scale
    assert not np.array_equal(f32.astype(np.float64), f64)  # the solve itself ran in float32


def _smooth3d(n, gamma=1.4):
    X, Y, Z = O.fv_uniform_meshgen(n)
    w = np.stack([1 + 0.3 * np.sin(2 * np.pi * (X + 2 * Y)) * np.cos(2 * np.pi * Z), 0.4 + 0.1 * np.cos(2 * np.pi * Y),
                  -0.3 + 0.2 * X, 0.2 * np.sin(2 * np.pi * Z), 1 + 0.2 * X * Y + 0.1 * Z])
    return O.compute_conservatives(w, gamma)


@pytest.mark.parametrize("n,p,rs", [((10, 9, 8), (3, 3, 3), "hllc"), ((20, 18, 17), (7, 7, 7), "llf"),
                                    ((24, 20, 1), (3, 3, 0), "hllc")], ids=["3d-p3-hllc", "3d-p7-llf", "2d-p3-hllc"])
def test_float32_stage_close_to_float64_stage(n, p, rs):
    u = _smooth3d(n)
    bc = (("outflow", "reflective"), ("periodic", "periodic"), ("free", "outflow"))
    kw = dict(n=n, p=p, gamma=1.4, riemann_solver=rs, bc=bc)
    dt64, d64 = O.hydrodynamics(O.Scheme(**kw), u)
    dt32, d32 = O.hydrodynamics(O.Scheme(bulk="fp32", **kw), u)
    assert d32.dtype == np.float64
    assert abs(dt32 - dt64) < 1e-6 * dt64
    scale = np.abs(d64).max()
    assert np.abs(d32 - d64).max() < 2e-4 * scale  # float32 rounding of O(1) values divided by h
    assert not np.array_equal(d32, d64)
