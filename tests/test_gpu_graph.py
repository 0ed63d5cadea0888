"""Graph-captured steps (ODE.advance(graph=True)) against eager steps: every
replay, interleaved with eager steps and host synchronisations, must give the
eager trajectory bit for bit (state, time, trouble log, predictor counters)."""

import numpy as np
import pytest

from fvhoe_amd import _lib
from fvhoe_amd import initial_conditions as ics
from fvhoe_amd.predictor import TroubledCellPredictor
from gpu_helpers import assert_bits_equal, solver_from

pytestmark = pytest.mark.gpu

P, O = ("periodic", "periodic"), ("outflow", "outflow")


def _pair(kind, integ, **kw):
    n = 64
    u0 = ics.kelvin_helmholtz(n, seed=0) if kind == "kh" else ics.four_quadrants(n, seed=2)
    bc = (P, P, P) if kind == "kh" else (O, O, P)
    mk = lambda: solver_from(u0, (n, n, 1), (3, 3, 0), bc, riemann_solver="hllc",  # noqa: E731
                             a_posteriori_slope_limiting=True, **kw)
    a, b = mk(), mk()
    a.integrator = b.integrator = integ
    return a, b


@pytest.mark.parametrize("kind,integ,form", [("kh", "ssprk3", "fp64"), ("kh", "rk4", "fp64"),
                                             ("quad", "ssprk3", "fp32"), ("quad", "ssprk3", "learned"),
                                             ("quad", "ssprk3", "learned-overlap"), ("quad", "rk4", "learned-overlap")])
def test_graph_replays_equal_eager_steps(kind, integ, form):
    kw = {"bulk_precision": "fp64" if form == "fp64" else "fp32"}
    if form.startswith("learned"):
        kw["predictor"] = TroubledCellPredictor.load()
        kw["predictor_overlap"] = form == "learned-overlap"
    a, b = _pair(kind, integ, **kw)
    torch = _lib.torch()

    def check(tag):
        torch.cuda.synchronize()
        assert_bits_equal(b.u.cpu().numpy(), a.u.cpu().numpy(), f"u after {tag}")
        assert a.t == b.t, tag
        assert a.step_count == b.step_count
        assert a.f_evaluation_count == b.f_evaluation_count
        assert_bits_equal(b.log["mean trouble"].cpu().numpy(), a.log["mean trouble"].cpu().numpy(), "log")
        if form.startswith("learned"):
            assert np.array_equal(a.predictor_stats.cpu().numpy(), b.predictor_stats.cpu().numpy())

    a.advance(3)
    b.advance(3, graph=True)  # 1 eager step, capture, 2 replays
    check("3 steps")
    for k in range(3):  # replays one by one, with eager steps in between
        a.advance(1)
        b.advance(1, graph=True)
        check(f"replay {k}")
        a.advance(1)
        b.advance(1)
        check(f"eager {k}")
    a.advance(4)
    b.advance(4, graph=True)
    check("4 more replays")


def test_capture_keeps_the_cyclic_collector_off():
    """A dead solver in a reference cycle (with its own captured graph) must not be finalised while
    another capture is open; the capture guard collects before the capture and keeps the collector
    off during it."""
    import gc

    dead, _ = _pair("kh", "ssprk3")
    dead.advance(2, graph=True)  # holds a captured graph
    del dead
    a, _ = _pair("kh", "ssprk3")
    seen = []
    helper = a.step_helper_function
    a.step_helper_function = lambda: (seen.append(gc.isenabled()), helper())
    a.advance(3, graph=True)  # 1 eager step, capture, 2 replays
    assert seen == [True, False]  # the eager step with the collector on, the captured one with it off
    assert gc.isenabled()


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_composed_graph_replays_equal_eager_steps(form):
    """The composed 3-D stage (divergence and RK combination in one kernel) captured in a graph."""
    n = (16, 12, 10)
    rng = np.random.default_rng(4)
    x = np.linspace(0, 1, n[0])[:, None, None]
    w = np.empty((5, *n))
    w[0] = 1.0 + 0.2 * np.sin(2 * np.pi * x) + 0.05 * rng.random(n)
    w[1], w[2], w[3] = 0.3, -0.2, 0.1
    w[4] = 1.0 + 0.1 * rng.random(n)
    from oracle import fvhoe_oracle as O

    u0 = O.compute_conservatives(w, 1.4)
    mk = lambda: solver_from(u0, n, (3, 3, 3), (P, P, P), riemann_solver="llf",  # noqa: E731
                             a_posteriori_slope_limiting=False, bulk_precision=form)
    a, b = mk(), mk()
    assert not a.fused
    a.integrator = b.integrator = "ssprk3"
    a.advance(4)
    b.advance(4, graph=True)  # 1 eager step, capture, 3 replays
    _lib.torch().cuda.synchronize()
    assert a.t == b.t and a.step_count == b.step_count
    assert_bits_equal(b.u.cpu().numpy(), a.u.cpu().numpy(), f"composed {form}")
