"""The IC classes, fv_average and the conservative conversion on the host
against what the reference solver builds (tests/golden/ics.npz, written by
tests/golden/make_golden.py from fvhoe/initial_conditions.py, fvhoe/fv.py:227-277
and fvhoe/solver.py:255-306), bit for bit."""

import warnings

import numpy as np
import pytest

from cases import IC_CASES
from fvhoe_amd import initial_conditions as ics
from helpers import load

GOLD = load("ics.npz")
CASES = sorted({k.split("/")[0] for k in GOLD})


@pytest.mark.parametrize("case", IC_CASES, ids=[c[0] for c in IC_CASES])
def test_initial_conditions_match_reference(case):
    key, cls, kw, n, p, dom = case
    ic = getattr(ics, cls)(**kw)
    X, Y, Z = ics.uniform_mesh(n, *dom)
    h = tuple((b - a) / k for (a, b), k in zip(dom, n))
    w0 = ic.build_ic()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)  # Sedov's mean spacing of a 1-cell axis is NaN there too
        w = w0(X, Y, Z)
        u0 = ics.fv_average(lambda x, y, z: ics.primitives_to_conservatives(w0(x, y, z), 1.4), X, Y, Z, h, p)
    assert np.array_equal(w.view(np.int64), GOLD[f"{key}/w"].view(np.int64)), key
    assert np.array_equal(u0.view(np.int64), GOLD[f"{key}/u0"].view(np.int64)), key


def test_every_case_is_in_the_fixture():
    assert sorted(c[0] for c in IC_CASES) == CASES
