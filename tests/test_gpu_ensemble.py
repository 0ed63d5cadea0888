"""The batched ensemble of 1-D runs against single runs of the fused stage
(themselves bit-identical to the reference) and the golden Shu-Osher
trajectory of the oracle."""

import numpy as np
import pytest

from fvhoe_amd import _lib
from fvhoe_amd import initial_conditions as ics
from fvhoe_amd.ensemble import Ensemble
from gpu_helpers import assert_bits_equal, solver_from
from helpers import group, load

pytestmark = pytest.mark.gpu

TRAJ = load("trajectories.npz")
P, O = ("periodic", "periodic"), ("outflow", "outflow")


def test_single_instance_matches_golden_shu_osher():
    g = group(TRAJ, "shuosher")
    e = Ensemble([g["u0"]], x=(-5, 5), p=5, riemann_solver="hllc", tile=128)
    for _ in range(6):
        e.step(1e9)
    assert_bits_equal(e.state(0).reshape(g["u_final"].shape), g["u_final"], "u")
    assert e.times()[0] == g["timestamps"][-1]


def _instances(ns, seed=3):
    rng = np.random.default_rng(seed)
    return [ics.shu_osher_instance(n, float(rng.uniform(0.1, 0.3)), float(rng.uniform(3, 7))) for n in ns]


@pytest.mark.parametrize("tile", [64, 128, 256])
def test_uneven_instances_match_single_runs(tile):
    ns = [64, 97, 128, 203]
    u0 = _instances(ns)
    T = 0.08
    e = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=tile)
    steps = e.run(T, graph=False, check_every=8)
    assert np.all(e.times() == T)
    for i, n in enumerate(ns):
        s = solver_from(u0[i], (n, 1, 1), (5, 0, 0), (O, P, P), x=(-5, 5), riemann_solver="hllc",
                        a_posteriori_slope_limiting=True)
        assert s.fused
        s.ssprk3(T)
        assert_bits_equal(e.state(i), s.u.cpu().numpy().reshape(5, n), f"instance {i} (n={n})")
        assert s.t == e.times()[i]
    print(f"[ensemble tile={tile}] {steps} ensemble steps", flush=True)


def test_graph_replay_equals_eager_and_keeps_finished_instances():
    ns = [48, 160, 256]
    u0 = _instances(ns, seed=5)
    T = 0.05
    a = Ensemble(u0, x=(-5, 5), p=3, riemann_solver="hllc", tile=64)
    b = Ensemble(u0, x=(-5, 5), p=3, riemann_solver="hllc", tile=64)
    a.run(T, graph=False, check_every=4)
    b.run(T, graph=True, check_every=4)
    for i in range(len(ns)):
        assert_bits_equal(b.state(i), a.state(i), f"instance {i}")
    assert np.array_equal(a.times(), b.times())
    # further replays past the target leave every instance as it is
    before = b.u.clone()
    b.run(T, graph=True, check_every=3)
    assert bool((b.u == before).all().item())


def test_fp32_ensemble_l1():
    ns = [128, 256]
    u0 = _instances(ns, seed=7)
    T = 0.1
    a = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc")
    b = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", bulk_precision="fp32")
    a.run(T)
    b.run(T)
    for i in range(len(ns)):
        l1 = float(np.mean(np.abs(a.state(i)[0] - b.state(i)[0])))
        print(f"[ensemble fp32] n={ns[i]} L1(rho) vs fp64 = {l1:.2e}", flush=True)
        assert l1 < 1e-5


def test_autotune_keeps_results():
    ns = [100, 300]
    u0 = _instances(ns, seed=9)
    T = 0.03
    ref = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=128)
    ref.run(T, graph=False)
    e = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=128)
    timings = e.autotune(T, steps=3)
    assert {k[0] for k in timings} == {64, 128, 256}
    assert (e.tile, "persistent" if e.persistent else "chain") in timings
    e.run(T)
    for i in range(len(ns)):
        assert_bits_equal(e.state(i), ref.state(i), f"instance {i}")


@pytest.mark.parametrize("tile", [64, 256])
@pytest.mark.parametrize("bulk", ["fp64", "fp32"])
def test_persistent_stage_equals_chain(tile, bulk):
    """One persistent kernel per RK stage (grid barriers between its phases) against the chain of launches,
    bit for bit, eagerly and as graph replays, past the target time too."""
    ns = [64, 97, 300, 203, 1200]
    u0 = _instances(ns, seed=11)
    T = 0.06
    a = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=tile, bulk_precision=bulk)
    b = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=tile, bulk_precision=bulk, persistent=True)
    c = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=tile, bulk_precision=bulk, persistent=True)
    sa = a.run(T, graph=False, check_every=4)
    sb = b.run(T, graph=False, check_every=4)
    sc = c.run(T, graph=True, check_every=4)
    print(f"[ensemble persistent tile={tile} {bulk}] steps {sa} {sb} {sc}", flush=True)
    for i in range(len(ns)):
        assert_bits_equal(b.state(i), a.state(i), f"eager instance {i}")
        assert_bits_equal(c.state(i), a.state(i), f"graph instance {i}")
    assert np.array_equal(a.times(), b.times()) and np.array_equal(a.times(), c.times())
    assert np.array_equal(a.steps.cpu().numpy(), c.steps.cpu().numpy())


def test_autotune_picks_a_stage_form():
    e = Ensemble(_instances([256, 512, 1024], seed=12), x=(-5, 5), p=5, riemann_solver="hllc")
    res = e.autotune(0.05, steps=3)
    assert set(res) == {(t, f) for t in (64, 128, 256) for f in ("chain", "persistent")}
    best = min(res, key=res.get)
    assert (e.tile, "persistent" if e.persistent else "chain") == best


@pytest.mark.parametrize("kinds", [("ic", "ic"), ("outflow", "ic"), ("dirichlet", "free")], ids=["ic-ic", "out-ic",
                                                                                                   "dir-free"])
def test_slab_boundaries_match_single_runs(kinds):
    """"ic" and constant "dirichlet" sides (the reference runs Shu-Osher with "ic" boundaries,
    tests/test_boundary_conditions.py:146-224): every instance equals its single fused run with the same
    boundaries, bit for bit."""
    from fvhoe_amd.boundary_conditions import BoundaryCondition
    from fvhoe_amd.solver import EulerSolver

    c = np.array([3.5, 8.0, 0.0, 0.0, 30.0])
    vals = tuple(c if k == "dirichlet" else None for k in kinds)
    rng = np.random.default_rng(14)
    ns = [80, 257, 640]
    solvers = []
    for n in ns:
        w0 = ics.ShuOsherInstance(eps=float(rng.uniform(0.1, 0.3)), k=float(rng.uniform(3, 7)))
        s = EulerSolver(w0=w0, nx=n, px=5, x=(-5, 5), gamma=1.4, riemann_solver="hllc",
                        a_posteriori_slope_limiting=True, progress_bar=False,
                        bc=BoundaryCondition(x=kinds, x_value=vals))
        assert s.fused
        solvers.append(s)
    T = 0.05
    e = Ensemble([s.u.cpu().numpy() for s in solvers], x=(-5, 5), p=5, riemann_solver="hllc", tile=64,
                 bc=BoundaryCondition(x=kinds, x_value=vals), ic=[s.u0 for s in solvers])
    e.run(T, graph=True, check_every=4)
    for i, s in enumerate(solvers):
        s.ssprk3(T)
        print(f"[ensemble {kinds}] instance {i} n={ns[i]} t={s.t!r} steps={s.step_count}", flush=True)
        assert s.t == e.times()[i]
        assert_bits_equal(e.state(i), s.u.cpu().numpy().reshape(5, ns[i]), f"instance {i}")


def test_forced_trouble_long_runs_match_single_runs():
    """Every cell flagged (pure MUSCL, fvhoe/solver.py:705-706) on runs long enough that an instance's
    fallback list takes several passes of its blocks, beside a short run: bit-identical to single fused runs
    with force_trouble."""
    ns = [3000, 64, 1500]
    u0 = _instances(ns, seed=21)
    T = 0.02
    e = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=256)
    e._cfg.force_trouble = 1
    steps = e.run(T, graph=True, check_every=4)
    assert np.all(e.times() == T)
    for i, n in enumerate(ns):
        s = solver_from(u0[i], (n, 1, 1), (5, 0, 0), (O, P, P), x=(-5, 5), riemann_solver="hllc",
                        a_posteriori_slope_limiting=True, force_trouble=True)
        assert s.fused
        s.ssprk3(T)
        assert_bits_equal(e.state(i), s.u.cpu().numpy().reshape(5, n), f"instance {i} (n={n})")
        assert s.t == e.times()[i]
    print(f"[ensemble forced trouble] {steps} ensemble steps", flush=True)
