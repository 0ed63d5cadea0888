"""Host logic of fvhoe_amd.fv.sweep_chain (no GPU): which consecutive sweeps run as one pair kernel.
The pair kernels themselves are checked bit for bit against single sweeps in tests/test_gpu_sweeps.py."""

import pytest

from fvhoe_amd import _lib
from fvhoe_amd import fv


class _Arr:
    def __init__(self, shape):
        self.shape = tuple(shape)

    def dim(self):
        return len(self.shape)


@pytest.fixture
def calls(monkeypatch):
    log = []
    monkeypatch.setattr(_lib, "as_device_bulk", lambda u: (u, False))
    monkeypatch.setattr(_lib, "back", lambda x, was_np: x)

    def zpair(x, kind, p, axis, zfirst):
        log.append(("zpair", p, axis, zfirst))
        return x

    def xypair(x, kind, p, first):
        log.append(("xypair", p, first))
        return x

    def single(x, mode, p, axis, pos=None):
        log.append(("single", p, axis))
        return x

    monkeypatch.setattr(fv, "_zpair", zpair)
    monkeypatch.setattr(fv, "_xypair", xypair)
    monkeypatch.setattr(fv, "stencil_sweep", single)
    return log


def test_cell_centres_pair_y_with_z(calls):
    fv.sweep_chain(_Arr((5, 40, 40, 80)), "c", ((3, 1), (3, 2), (3, 3)))
    assert calls == [("single", 3, 1), ("zpair", 3, 2, False)]


@pytest.mark.parametrize("steps,want", [
    (((7, 2), (7, 1)), [("xypair", 7, 2)]),                       # z fluxes' c chain: y then x
    (((7, 1), (7, 2)), [("xypair", 7, 1)]),                       # z fluxes' t chain: x then y
    (((7, 3), (7, 2)), [("zpair", 7, 2, True)]),                  # x fluxes' c chain: z then y
    (((7, 1), (7, 3)), [("zpair", 7, 1, False)]),                 # y fluxes' chains: x then z
    (((3, 1), (2, 2)), [("single", 3, 1), ("single", 2, 2)]),     # different degrees: no pair
    (((0, 1), (0, 2)), [("single", 0, 1), ("single", 0, 2)]),     # degree 0: no pair
])
def test_pairs(calls, steps, want):
    fv.sweep_chain(_Arr((5, 40, 40, 80)), "t", steps)
    assert calls == want
