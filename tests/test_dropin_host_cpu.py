"""Drop-in functions that are host arithmetic or index metadata, with no
device work: advection_dt (fvhoe/hydro.py:249-267), get_symmetric_slices
(fvhoe/fv.py:88-103), get_legendre_quadrature (fvhoe/fv.py:227-240)."""

import numpy as np
import pytest

import oracle as O
from oracle import fvhoe_oracle as OF
from fvhoe_amd import fv, hydro

RNG = np.random.default_rng(18)


def test_host_helpers():
    assert hydro.advection_dt(0.01, 1.0, 1.0, 1.0, CFL=0.8) == 0.8 * 0.01 / (1.0 + 1.0 + 1.0)
    assert hydro.advection_dt(0.5, vx=2.0) == 0.8 * 0.5 / (2.0 + 0.0 + 0.0)
    u = RNG.standard_normal((5, 9, 8, 7))
    for axis in (1, 2, 3):
        for k in (1, 3, 5):
            sl = fv.get_symmetric_slices(4, k, axis)
            assert len(sl) == k
            for i, s in enumerate(sl):
                assert np.array_equal(u[s], OF._shift(u, axis, i, k - 1 - i))
    with pytest.raises(ValueError):
        fv.get_symmetric_slices(4, 0, 1)
    pts, wts = fv.get_legendre_quadrature(5)
    assert len(pts) == 3 and abs(wts.sum() - 1.0) < 1e-15 and np.all(np.abs(pts) < 0.5)
