"""The large-grid sweep kernels against the oracle bit for bit: the
sliding-window sweeps (arrays of 64K elements and more), the z-paired sweeps
(z of 64 cells and more), and a 3-D composed stage on a grid that uses both
(fvhoe/fv.py:106-224, fvhoe/solver.py:429-651)."""

import numpy as np
import pytest

from fvhoe_amd import _lib
from fvhoe_amd import fv
from gpu_helpers import assert_bits_equal, solver_from
from oracle import fvhoe_oracle as O

pytestmark = pytest.mark.gpu


def _oracle_sweep(u, pos, p, axis):
    if pos == "t":
        return O.transverse_reconstruction(u, p, axis)
    return O.conservative_interpolation(u, p, axis, pos)


@pytest.mark.parametrize("pos", ["l", "c", "r", "t"])
@pytest.mark.parametrize("axis", [1, 2, 3])
def test_sliding_sweeps(pos, axis):
    rng = np.random.default_rng(10 * axis + "lcrt".index(pos))
    u = rng.standard_normal((5, 30, 26, 70))
    for p in range(0, 9):
        got = fv.stencil_sweep(u, "reconstruction" if pos == "t" else "interpolation", p, axis,
                               None if pos == "t" else pos)
        assert_bits_equal(got, _oracle_sweep(u, pos, p, axis), f"{pos} p={p} axis={axis}")


@pytest.mark.parametrize("pos", ["c", "t"])
@pytest.mark.parametrize("order", [((1,), 3), ((2,), 3), (3, (1,)), (3, (2,))], ids=["x-z", "y-z", "z-x", "z-y"])
def test_zpair_sweeps(pos, order):
    rng = np.random.default_rng(3)
    u = rng.standard_normal((5, 21, 19, 75))
    axes = [a if isinstance(a, int) else a[0] for a in order]
    for p in range(1, 9):
        x = _lib.as_device(u)[0]
        steps = tuple((p, a) for a in axes)
        got = fv.sweep_chain(x, pos, steps).cpu().numpy()
        pair = fv._zpair(x, _lib.STENCIL[pos], p, axes[0] if axes[1] == 3 else axes[1], axes[0] == 3)
        assert pair is not None  # the pair kernel ran
        want = _oracle_sweep(_oracle_sweep(u, pos, p, axes[0]), pos, p, axes[1])
        assert_bits_equal(got, want, f"{pos} p={p} axes={axes}")
        assert_bits_equal(pair.cpu().numpy(), want, f"{pos} p={p} axes={axes} (pair)")


@pytest.mark.parametrize("p,rs", [(3, "hllc"), (7, "llf"), (2, "advection_upwind")])
def test_composed_3d_stage_with_long_z(p, rs):
    n = (18, 17, 72)
    rng = np.random.default_rng(p)
    X = np.linspace(0, 1, n[2])[None, None, None, :]
    w = np.empty((5, *n))
    w[0] = 1.0 + 0.2 * np.sin(2 * np.pi * X[0]) + 0.05 * rng.random(n)
    w[1], w[2], w[3] = 0.3, -0.2, 0.5 + 0.1 * rng.random(n)
    w[4] = 1.0 + 0.1 * rng.random(n)
    u = O.compute_conservatives(w, 1.4)
    P = ("periodic", "periodic")
    s = solver_from(u, n, (p, p, p), (P, ("outflow", "outflow"), P), riemann_solver=rs, fused=False)
    dt, dudt = s.hydrodynamics(s.u)
    sch = O.Scheme(n=n, p=(p, p, p), gamma=1.4, riemann_solver=rs,
                   bc=(P, ("outflow", "outflow"), P))
    dto, dudto = O.hydrodynamics(sch, u)
    print(f"[3d p={p} {rs}] dt={dt!r}", flush=True)
    assert dt == dto
    assert_bits_equal(dudt.cpu().numpy(), dudto, "dudt")


@pytest.mark.parametrize("f32", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("axis", [1, 2])
def test_fused_fill_and_sweep(axis, f32):
    """fvhoe_bc_stencil_sweep == fill + sweep bit for bit, every non-slab kind on every side, and both
    equal the oracle's fill followed by its sweep (in float32 after the fill for the FP32 form)."""
    from fvhoe_amd.boundary_conditions import BoundaryCondition

    torch = _lib.torch()
    rng = np.random.default_rng(axis)
    u = rng.standard_normal((5, 23, 21, 40))
    x = _lib.as_device(u)[0]
    dtype = torch.float32 if f32 else None
    for kinds in ((("periodic", "periodic"),) * 3, (("outflow", "reflective"), ("free", "reflective"), ("reflective", "free"))):
        bc = BoundaryCondition(x=kinds[0], y=kinds[1], z=kinds[2])
        for p in (1, 3, 7, 8):
            g = (8, 7, 6)
            got = bc.apply_swept(x, g, "c", p, axis, dtype=dtype)
            assert got is not None  # the fused kernel ran
            ref = fv.stencil_sweep(bc.apply(x, g, dtype=dtype), "interpolation", p, axis, "c")
            assert got.dtype == ref.dtype and got.shape == ref.shape
            if f32:
                assert torch.equal(got.view(torch.int32), ref.view(torch.int32)), (kinds, p, axis)
            else:
                assert_bits_equal(got.cpu().numpy(), ref.cpu().numpy(), f"{kinds} p={p} axis={axis}")
            filled = O.apply_bc(u, g, kinds)
            want = O.conservative_interpolation(filled.astype(np.float32) if f32 else filled, p, axis, "c")
            assert np.array_equal(got.cpu().numpy().view(np.uint8), want.view(np.uint8)), (kinds, p, axis, "oracle")


@pytest.mark.parametrize("f32", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("pos", ["c", "t"])
@pytest.mark.parametrize("first", [1, 2], ids=["x-y", "y-x"])
def test_xypair_sweeps(first, pos, f32):
    """fvhoe_stencil_sweep_xypair == two single sweeps == the oracle's two sweeps, bit for bit (the
    oracle's float32 sweeps for the FP32 form)."""
    rng = np.random.default_rng(7 + first)
    u = rng.standard_normal((5, 29, 27, 33))
    x = _lib.as_device(u)[0]
    if f32:
        x = x.float()
    mode = "reconstruction" if pos == "t" else "interpolation"
    for p in range(1, 9):
        pair = fv._xypair(x, _lib.STENCIL[pos], p, first)
        assert pair is not None  # the pair kernel ran
        ref = fv.stencil_sweep(fv.stencil_sweep(x, mode, p, first, None if pos == "t" else pos), mode, p, 3 - first,
                               None if pos == "t" else pos)
        assert pair.dtype == ref.dtype and pair.shape == ref.shape
        if f32:
            assert _lib.torch().equal(pair.view(_lib.torch().int32), ref.view(_lib.torch().int32)), (pos, p, first)
        else:
            assert_bits_equal(pair.cpu().numpy(), ref.cpu().numpy(), f"{pos} p={p} first={first}")
        uo = u.astype(np.float32) if f32 else u
        want = _oracle_sweep(_oracle_sweep(uo, pos, p, first), pos, p, 3 - first)
        assert np.array_equal(pair.cpu().numpy().view(np.uint8), want.view(np.uint8)), (pos, p, first, "oracle")
