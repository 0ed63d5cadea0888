"""The C-ABI library loads and exports every symbol include/fvhoe_hip.h declares;
host-only entry points agree with the oracle (no GPU needed)."""

import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from fvhoe_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "fvhoe_hip.h")).read()
    return sorted(set(re.findall(r"^int (fvhoe_\w+)\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.exported_symbols())
    assert lib.fvhoe_version() == 1


@pytest.mark.parametrize("p", range(9))
@pytest.mark.parametrize("kind", ["l", "c", "r", "t"])
def test_stencil_weights_match_exact_rationals(p, kind):
    out = (ctypes.c_double * 9)()
    n = ctypes.c_int(0)
    assert _lib.lib().fvhoe_stencil_weights(p, _lib.STENCIL[kind], out, ctypes.byref(n)) == 0
    got = np.array(out[: n.value])
    want = O.transverse_weights(p) if kind == "t" else O.interpolation_weights(p, kind)
    assert np.array_equal(got.view(np.int64), want.view(np.int64))


def test_stencil_weights_rejects_bad_degree():
    out = (ctypes.c_double * 9)()
    n = ctypes.c_int(0)
    assert _lib.lib().fvhoe_stencil_weights(9, 0, out, ctypes.byref(n)) == _lib.ERR_ARG


def _cfg(n, p, apost=True):
    c = _lib.FvhoeStageConfig()
    c.n[:] = list(n)
    c.p[:] = list(p)
    c.h[:] = [1.0 / max(k, 1) for k in n]
    c.gamma, c.cfl, c.csq_floor = 1.4, 0.8, 1e-10
    c.riemann = 2
    c.a_posteriori = int(apost)
    c.limiter = 1
    c.nad_eps = 1e-2
    c.nad_vars = 31
    c.pad_var = 4
    c.pad_hi = float("inf")
    c.sed = 1
    c.sed_eps = 1e-10
    return c


@pytest.mark.parametrize("n,p,ok", [((2048, 2048, 1), (3, 3, 0), 1), ((256, 1, 1), (3, 0, 0), 1),
                                    ((64, 64, 1), (3, 2, 0), 0), ((16, 16, 16), (7, 7, 7), 0),
                                    ((1, 64, 1), (0, 3, 0), 0), ((6, 6, 1), (3, 3, 0), 0)])
def test_stage_support_matrix(n, p, ok):
    assert _lib.lib().fvhoe_stage_supported(_cfg(n, p)) == ok


@pytest.mark.parametrize("fp32", [0, 1])
def test_stage_workspace_size(fp32):
    c = _cfg((2048, 2048, 1), (3, 3, 0))
    c.bulk_fp32 = fp32
    b = ctypes.c_ulonglong(0)
    assert _lib.lib().fvhoe_stage_workspace_bytes(ctypes.byref(c), ctypes.byref(b)) == 0
    n = 2048
    if fp32:
        # per-block reduction partials of the FP32 column-strip prims kernel: 60 valid columns per
        # wave (64 lanes - 2 x (c + t reach 1 + 1)), 4 waves of 4 rows per block
        blocks = ((n + 8 + 59) // 60) * ((n + 8 + 15) // 16)
    else:
        # FP64 prims: one block per 32 x 32 tile of the (n + 8)^2 W box
        blocks = ((n + 8 + 31) // 32) ** 2
    esz = 4 if fp32 else 8
    # W, F, G, WS; flags and the fallback list; the partials; the list of cells the fused candidate check
    # leaves to the SED pass with their NAD indicators (a-posteriori with the SED on)
    expect = esz * 5 * ((n + 8) ** 2 + 2 * (n + 1) * n + (n + 6) ** 2) + 5 * n * n + 88 * blocks
    expect += (4 + 8) * n * n if c.sed else 0
    assert expect <= b.value <= expect + 4096


def test_stage_rejects_null_pointers():
    c = _cfg((64, 64, 1), (3, 3, 0))
    assert _lib.lib().fvhoe_stage(ctypes.byref(c), None, None, None, None, None, None) == _lib.ERR_ARG
