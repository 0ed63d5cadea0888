"""fvhoe_stage_rk_static, the stage entry point that allocates nothing (its
scratch is a fixed device array inside the library), against fvhoe_stage_rk
with a caller workspace, through the C-ABI, bit for bit: two calls in a row,
the second on a smaller grid, in FP64 and FP32, 2-D and 1-D; and the clear
error for a grid whose scratch does not fit."""

import ctypes

import numpy as np
import pytest

from fvhoe_amd import _lib
from fvhoe_amd import initial_conditions as ics
from gpu_helpers import assert_bits_equal, solver_from

pytestmark = pytest.mark.gpu
P, O = ("periodic", "periodic"), ("outflow", "outflow")


def _solver(n, two_d, p, bulk):
    if two_d:
        return solver_from(ics.kelvin_helmholtz(n, seed=4), (n, n, 1), (p, p, 0), (P, P, P), riemann_solver="hllc",
                           a_posteriori_slope_limiting=True, bulk_precision=bulk)
    return solver_from(ics.sod_shock_tube(n, p), (n, 1, 1), (p, 0, 0), (O, P, P), riemann_solver="hllc",
                       a_posteriori_slope_limiting=True, bulk_precision=bulk)


def _run(s, static):
    """One fused SSP-RK3 first stage (AXPY combine, dt clipped to a target) from the solver's state."""
    t = _lib.torch()
    u = s.u
    cfg = _lib.FvhoeStageConfig.from_buffer_copy(s._cfg)
    cfg.mag_count = cfg.mag_idx = cfg.mag_nad = cfg.mag_pad = None
    dudt, out = _lib.empty(u.shape), _lib.empty(u.shape)
    dt, dt_step, t_dev = _lib.zeros((1,)), _lib.zeros((1,)), _lib.zeros((1,))
    trouble = _lib.empty(tuple(s.n), dtype=t.int8)
    rk = _lib.FvhoeRkFuse()
    rk.kind = _lib.RK_AXPY
    rk.out, rk.u0, rk.dt_step = out.data_ptr(), u.data_ptr(), dt_step.data_ptr()
    rk.set_dt_step, rk.t, rk.target, rk.has_target = 1, t_dev.data_ptr(), 1e-3, 1
    if static:
        _lib.call("fvhoe_stage_rk_static", ctypes.byref(cfg), _lib.ptr(u), _lib.ptr(dudt), _lib.ptr(dt),
                  _lib.ptr(trouble), ctypes.byref(rk), _lib.stream())
    else:
        _lib.call("fvhoe_stage_rk", ctypes.byref(cfg), _lib.ptr(u), _lib.ptr(dudt), _lib.ptr(dt), _lib.ptr(trouble),
                  s._workspace(), ctypes.byref(rk), _lib.stream())
    t.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in dict(dudt=dudt, out=out, dt=dt, dt_step=dt_step, trouble=trouble).items()}


@pytest.mark.parametrize("bulk", ["fp64", "fp32"])
@pytest.mark.parametrize("two_d,sizes,p", [(True, (256, 96), 3), (True, (160, 64), 5), (False, (4096, 600), 5)],
                         ids=["2d-p3", "2d-p5", "1d-p5"])
def test_static_entry_equals_workspace_entry(two_d, sizes, p, bulk):
    for n in sizes:  # the second call on a smaller grid: the scratch keeps the first call's contents
        s = _solver(n, two_d, p, bulk)
        need = ctypes.c_ulonglong(0)
        _lib.call("fvhoe_stage_workspace_bytes", ctypes.byref(s._cfg), ctypes.byref(need))
        assert need.value <= _lib.STATIC_SCRATCH_BYTES, (n, need.value)
        print(f"[static {two_d and '2d' or '1d'} n={n} p={p} {bulk}] scratch {need.value} bytes", flush=True)
        a = _run(s, static=False)
        b = _run(s, static=True)
        for k in a:
            if k == "trouble":
                assert np.array_equal(a[k], b[k]), (n, k)
            else:
                assert_bits_equal(b[k], a[k], f"n={n} {k}")
        assert int(a["trouble"].sum()) > 0  # the check and the fallback ran


def test_static_entry_rejects_a_grid_beyond_its_scratch():
    s = _solver(1024, True, 3, "fp64")
    need = ctypes.c_ulonglong(0)
    _lib.call("fvhoe_stage_workspace_bytes", ctypes.byref(s._cfg), ctypes.byref(need))
    assert need.value > _lib.STATIC_SCRATCH_BYTES
    with pytest.raises(ValueError, match="fixed device scratch"):
        _run(s, static=True)
