"""integrate() with the steps between snapshots replayed in chunks (one host
synchronisation per chunk) against integrate() stepping with a host
synchronisation per step (fvhoe/ode.py:94-193): the same snapshots,
timestamps, step and RHS counts, trouble log and predictor counters, bit for
bit, also when every chunk overshoots its target by many no-op steps."""

import numpy as np
import pytest

from fvhoe_amd import initial_conditions as ics
from fvhoe_amd.boundary_conditions import BoundaryCondition
from fvhoe_amd.predictor import TroubledCellPredictor
from fvhoe_amd.solver import EulerSolver
from gpu_helpers import assert_bits_equal, solver_from

pytestmark = pytest.mark.gpu

P, O = ("periodic", "periodic"), ("outflow", "outflow")


def _kh(**kw):
    n = 64
    return solver_from(ics.kelvin_helmholtz(n, seed=1), (n, n, 1), (3, 3, 0), (P, P, P), riemann_solver="hllc",
                       a_posteriori_slope_limiting=True, **kw)


def _sod(**kw):
    return EulerSolver(w0=ics.Shock1D(), nx=128, px=3, gamma=1.4, bc=BoundaryCondition(x="outflow"),
                       riemann_solver="hllc", a_posteriori_slope_limiting=True, progress_bar=False, **kw)


def _quad(**kw):
    n = 64
    return solver_from(ics.four_quadrants(n, seed=3), (n, n, 1), (3, 3, 0), (O, O, P), riemann_solver="hllc",
                       a_posteriori_slope_limiting=True, bulk_precision="fp32",
                       predictor=TroubledCellPredictor.load(), **kw)


def _compare(a, b, tag):
    assert b._chunkable() and not a._chunkable(), tag
    assert a.t == b.t, tag
    assert a.step_count == b.step_count, tag
    assert a.f_evaluation_count == b.f_evaluation_count, tag
    assert a.timestamps == b.timestamps, tag
    assert a.snapshot_times == b.snapshot_times, tag
    assert len(a.snapshots) == len(b.snapshots), tag
    for k, (sa, sb) in enumerate(zip(a.snapshots, b.snapshots)):
        assert_bits_equal(sb["w"], sa["w"], f"{tag} snapshot {k}")
        assert_bits_equal(sb["trouble"], sa["trouble"], f"{tag} snapshot {k} trouble")
        for key in ("NAD violation magnitude", "PAD violation magnitude"):
            assert_bits_equal(sb[key], sa[key], f"{tag} snapshot {k} {key}")
    assert_bits_equal(b.u.cpu().numpy(), a.u.cpu().numpy(), f"{tag} u")
    if a.predictor_stats is not None:
        assert np.array_equal(a.predictor_stats.cpu().numpy(), b.predictor_stats.cpu().numpy()), tag


CASES = [
    ("kh-ssprk3", _kh, "ssprk3", dict(T=0.06, downbeats=[0.02, 0.045])),
    ("sod-rk4", _sod, "rk4", dict(T=0.1, downbeats=np.array([0.03, 0.05]))),
    ("quad-learned", _quad, "ssprk3", dict(T=0.04, downbeats=[0.01])),
    ("kh-n", _kh, "ssprk2", dict(n=7)),
]


@pytest.mark.parametrize("launch", ["graph", "eager"])
@pytest.mark.parametrize("overshoot", [False, True], ids=["estimated", "overshoot"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_chunked_integrate_equals_per_step(case, overshoot, launch):
    name, mk, integ, kw = case
    a, b = mk(), mk()
    a._chunkable = lambda: False  # host-synchronised steps, as take_step does them
    b.CHUNK_GRAPHS = launch == "graph"
    if overshoot:  # every chunk launches 40 steps: the ones past each target must move nothing
        b._chunk_length = lambda target: 40
    getattr(a, integ)(**kw)
    getattr(b, integ)(**kw)
    print(f"[{name}] steps {a.step_count} t={a.t!r}", flush=True)
    _compare(a, b, name)
