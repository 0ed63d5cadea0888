"""The learned stage of the batched ensemble: the planar network (the 2-D
troubled-cell predictor on 1-D runs seen as 2-D states constant along y)
against the oracle's planar network on the oracle's features, and the
predicted step's semantics bit for bit (a never-flagging network gives the
plain step, an always-flagging one gives the step with every cell forced
troubled), graph replays equal to eager steps, and the trained network on a
small Shu-Osher ensemble with its counters."""

import numpy as np
import pytest

from fvhoe_amd import _lib
from fvhoe_amd import initial_conditions as ics
from fvhoe_amd.boundary_conditions import BoundaryCondition
from fvhoe_amd.ensemble import Ensemble, shu_osher_ensemble_solver
from fvhoe_amd.predictor import TroubledCellPredictor
from gpu_helpers import assert_bits_equal
from oracle import fvhoe_oracle as O
from oracle import predictor_oracle as PO

pytestmark = pytest.mark.gpu


def _instances(ns, seed=3):
    rng = np.random.default_rng(seed)
    return [ics.shu_osher_instance(n, float(rng.uniform(0.1, 0.3)), float(rng.uniform(3, 7))) for n in ns]


def _theta(seed, bias=0.3):
    th = PO.init_params(seed)
    th[-1] = bias  # some logits on both sides of 0
    return th


@pytest.mark.parametrize("bc", ["outflow", "periodic", "reflective"])
@pytest.mark.parametrize("tile", [64, 256])
def test_planar_logits_match_oracle(bc, tile):
    ns = [300, 513, 97, 700]  # several tiles per run, overhanging last tiles
    u0 = _instances(ns, seed=11)
    th = _theta(4)
    e = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=tile, bc=BoundaryCondition(x=bc),
                 predictor=TroubledCellPredictor(th))
    z = e.predictor_logits().cpu().numpy()
    for i, n in enumerate(ns):
        s = O.Scheme(n=(n, 1, 1), p=(5, 0, 0), gamma=1.4, a_posteriori=True,
                     bc=((bc, bc), ("periodic", "periodic"), ("periodic", "periodic")), domain=((-5, 5), (0, 1), (0, 1)))
        w3 = O.interpolate_primitives_from_conservatives(s, u0[i].reshape(5, n, 1, 1), s.p, (3, 0, 0))[..., 0, 0]
        X = PO.features_1d(w3[:, 3:-3], w3.min(axis=1), w3.max(axis=1), (bc, bc))
        zo = PO.forward_planar(th, X)
        zi = z[e.offsets[i]:e.offsets[i + 1]]
        err = float(np.abs(zi - zo).max())
        print(f"[planar logits bc={bc} tile={tile} n={n}] max|dz| {err:.2e} of {np.abs(zo).max():.2e}; "
              f"{int((zo > 0).sum())} positive", flush=True)
        assert err <= 2e-3 * float(np.abs(zo).max()) + 1e-4, (i, err)


def _pair(ns, thr, **kw):
    u0 = _instances(ns, seed=5)
    pred = TroubledCellPredictor(_theta(6), logit_threshold=thr)
    a = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=64, count_troubles=True, **kw)
    b = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=64, predictor=pred, **kw)
    return a, b


def test_never_flagging_network_gives_the_plain_step():
    ns = [160, 257, 96]
    a, b = _pair(ns, 1e30)
    plain = Ensemble(_instances(ns, seed=5), x=(-5, 5), p=5, riemann_solver="hllc", tile=64)
    for e in (a, b, plain):
        e.run(0.05, graph=False, check_every=4)
    for i in range(len(ns)):
        assert_bits_equal(b.state(i), plain.state(i), f"instance {i}")
        assert_bits_equal(a.state(i), plain.state(i), f"instance {i} (counting)")
    assert np.array_equal(b.times(), plain.times())
    sa, sb = a.predictor_stats.cpu().numpy(), b.predictor_stats.cpu().numpy()
    print(f"[never] plain counters {sa.tolist()} learned {sb.tolist()}", flush=True)
    assert sb[0] == 0 and sb[2] == 0 and sb[1] == sa[1] > 0


def test_always_flagging_network_gives_forced_trouble():
    ns = [160, 2600, 96]  # the long run's pre-flag list takes several passes of its fallback blocks
    u0 = _instances(ns, seed=5)
    forced = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=64)
    forced._cfg.force_trouble = 1
    always = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=64,
                      predictor=TroubledCellPredictor(_theta(2), logit_threshold=-1e30))
    for e in (forced, always):
        for _ in range(4):
            e.step(1.0)
    for i in range(len(ns)):
        assert_bits_equal(always.state(i), forced.state(i), f"instance {i}")
    st = always.predictor_stats.cpu().numpy()
    assert st[0] == 3 * 4 * sum(ns) and st[1] == 0  # every cell pre-flagged in every stage, none left


@pytest.mark.parametrize("bulk", ["fp64", "fp32"])
def test_learned_graph_replay_equals_eager(bulk):
    ns = [200, 333]
    u0 = _instances(ns, seed=8)
    pred = TroubledCellPredictor(_theta(3, bias=0.0), logit_threshold=0.0)
    a = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=128, predictor=pred, bulk_precision=bulk)
    b = Ensemble(u0, x=(-5, 5), p=5, riemann_solver="hllc", tile=128, predictor=pred, bulk_precision=bulk)
    a.run(0.04, graph=False, check_every=4)
    b.run(0.04, graph=True, check_every=4)
    for i in range(len(ns)):
        assert_bits_equal(b.state(i), a.state(i), f"instance {i}")
    assert np.array_equal(a.predictor_stats.cpu().numpy(), b.predictor_stats.cpu().numpy())
    assert int(a.predictor_stats[0].item()) > 0  # the network pre-flagged cells


# The trained network on a small Shu-Osher ensemble ("ic" boundaries, FP32 bulk form as the ensemble
# workload runs it): counters and the L1 density difference from the plain FP32 run (bound as the 2-D
# learned form's: 1e-2 of a density of order 1).
LEARNED_L1_BOUND = 1e-2


def test_trained_network_on_shu_osher_ensemble():
    kw = dict(count=12, seed=3, nmin=256, nmax=1024, bulk_precision="fp32", tile=128)
    plain = shu_osher_ensemble_solver(count_troubles=True, **kw)
    learned = shu_osher_ensemble_solver(predictor=TroubledCellPredictor.load(), **kw)
    T = 0.3
    plain.run(T, graph=True)
    learned.run(T, graph=True)
    assert np.all(plain.times() == T) and np.all(learned.times() == T)
    sp, sl = plain.predictor_stats.cpu().numpy(), learned.predictor_stats.cpu().numpy()
    l1 = max(float(np.abs(learned.state(i)[0] - plain.state(i)[0]).mean()) for i in range(plain.count))
    share = sl[1] / max(sp[1], 1)
    print(f"[shu-osher learned] check flags plain {sp[1]}, learned: pre-flagged {sl[0]}, second pass {sl[1]} "
          f"(share {share:.3f}), pre-flagged not flagged by the check {sl[2]}; max L1(rho) vs plain {l1:.2e}",
          flush=True)
    assert sp[1] > 0
    assert l1 <= LEARNED_L1_BOUND
