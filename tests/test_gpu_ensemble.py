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
    assert set(timings) == {64, 128, 256} and e.tile in timings
    e.run(T)
    for i in range(len(ns)):
        assert_bits_equal(e.state(i), ref.state(i), f"instance {i}")
