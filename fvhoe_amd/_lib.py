"""ctypes binding of libfvhoe_hip.so (the C-ABI in include/fvhoe_hip.h).

There is no fallback: if the library is missing, or no GPU is visible when a
kernel is called, this module raises.  torch provides device memory and the
current stream only.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_longlong, c_ulonglong, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfvhoe_hip.so")

OK, ERR_ARG, ERR_UNSUPPORTED, ERR_LAUNCH, ERR_TOO_LARGE = 0, 1, 2, 3, 4
STATIC_SCRATCH_BYTES = 64 << 20  # FVHOE_STATIC_SCRATCH_BYTES: the fixed device scratch of fvhoe_stage_rk_static
STENCIL = {"l": 0, "c": 1, "r": 2, "t": 3}
RIEMANN = {"advection_upwind": 0, "llf": 1, "hllc": 2}
LIMITER = {None: 0, "minmod": 1, "moncen": 2}
BC = {"periodic": 0, "outflow": 1, "reflective": 2, "free": 3, "dirichlet": 4, "ic": 4}
RK_AXPY, RK_SSPRK3_STAGE3, RK_SSPRK3_FINAL, RK_SSPRK2_FINAL, RK_HALF_AXPY, RK_RK4_FINAL = range(6)
RK_SSPRK3_STAGE3_SUM, RK_SSPRK3_FINAL_SUM = 6, 7  # fused stage only


class FvhoeStageConfig(ctypes.Structure):
    _fields_ = [
        ("n", c_int * 3), ("p", c_int * 3), ("h", c_double * 3),
        ("gamma", c_double), ("cfl", c_double), ("csq_floor", c_double),
        ("bc", c_int * 6), ("riemann", c_int), ("a_posteriori", c_int), ("limiter", c_int),
        ("force_trouble", c_int), ("nad_eps", c_double), ("nad_mode", c_int), ("nad_range", c_int),
        ("nad_vars", c_int), ("pad_var", c_int), ("pad_lo", c_double), ("pad_hi", c_double),
        ("sed", c_int), ("sed_eps", c_double), ("fixed_dt_on", c_int), ("fixed_dt", c_double),
        ("bulk_fp32", c_int),
        ("slab", c_void_p * 6), ("slab_g", c_int), ("slab_pad", c_int * 3),
        ("mag_count", c_void_p), ("mag_idx", c_void_p), ("mag_nad", c_void_p), ("mag_pad", c_void_p),
        ("mag_slot", c_int),
    ]


class FvhoeRkFuse(ctypes.Structure):
    _fields_ = [
        ("kind", c_int), ("out", c_void_p), ("u0", c_void_p), ("k0", c_void_p), ("k1", c_void_p),
        ("k2", c_void_p), ("dt_step", c_void_p), ("set_dt_step", c_int), ("t", c_void_p),
        ("target", c_double), ("has_target", c_int), ("target_ptr", c_void_p),
    ]


class FvhoePredict(ctypes.Structure):
    _fields_ = [("packed", c_void_p), ("logit_threshold", c_float), ("infer", c_int), ("stats", c_void_p),
                ("overlap", c_int), ("side_stream", c_void_p), ("ev_fork", c_void_p), ("ev_join", c_void_p)]


PRED_NPARAMS = 8433
PRED_PACKED_BYTES = 24576
PRED_PLANAR_BYTES = 4608
PRED_HALO = 6

_P = c_void_p
_SIGNATURES = {
    "fvhoe_version": (c_int, []),
    "fvhoe_stencil_weights": (c_int, [c_int, c_int, POINTER(c_double), POINTER(c_int)]),
    "fvhoe_stencil_sweep": (c_int, [_P, _P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, _P]),
    "fvhoe_compute_primitives": (c_int, [_P, _P, c_int, c_longlong, c_double, _P]),
    "fvhoe_compute_conservatives": (c_int, [_P, _P, c_int, c_longlong, c_double, _P]),
    "fvhoe_compute_sound_speed": (c_int, [_P, _P, c_longlong, c_double, c_double, _P]),
    "fvhoe_hydro_dt": (c_int, [_P, c_longlong, c_double, c_int, c_double, c_double, c_double, _P, _P, _P]),
    "fvhoe_riemann": (c_int, [_P, _P, _P, c_int, c_longlong, c_int, c_int, c_double, c_double, _P]),
    "fvhoe_riemann_faces": (c_int, [_P, _P, _P, c_int, c_int, c_int, c_int, c_int, c_double, c_double, _P]),
    "fvhoe_muscl": (c_int, [_P, _P, _P, c_int, c_int, c_int, c_int, c_int, c_int, _P]),
    "fvhoe_face_fluxes": (c_int, [_P, _P, c_int, c_int, c_int, c_int, c_int, c_int, c_double, c_double, _P]),
    "fvhoe_stencil_sweep_zpair": (c_int, [_P, _P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, _P]),
    "fvhoe_bc_apply": (c_int, [_P, _P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, POINTER(c_int), _P]),
    "fvhoe_bc_apply_dirichlet": (c_int, [_P, _P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, POINTER(c_int),
                                         POINTER(c_void_p), _P]),
    "fvhoe_stage_ghost_depth": (c_int, [POINTER(FvhoeStageConfig), POINTER(c_int)]),
    "fvhoe_euler_equation": (c_int, [_P, _P, _P, _P, c_int, c_int, c_int, c_int, c_int, c_double, c_double,
                                     c_double, _P]),
    "fvhoe_minmax": (c_int, [_P, c_int, c_longlong, _P, _P, _P]),
    "fvhoe_detect_troubled_cells": (c_int, [_P, _P, c_int, c_int, c_int, c_int, _P, c_int, c_int, c_double,
                                            c_int, c_int, c_double, c_double, c_int, c_double, _P, _P, _P, _P]),
    "fvhoe_smooth_extrema_detector": (c_int, [_P, _P, c_int, c_int, c_int, c_int, c_int, c_double, _P]),
    "fvhoe_broadcast_simple": (c_int, [_P, _P, _P, _P, c_int, c_int, c_int, c_int, _P]),
    "fvhoe_shift_max": (c_int, [_P, POINTER(c_longlong), c_longlong, _P, POINTER(c_longlong), c_longlong,
                                c_double, c_int, _P, POINTER(c_longlong), c_longlong, POINTER(c_int), _P]),
    "fvhoe_blend": (c_int, [_P, _P, _P, c_int, c_longlong, _P]),
    "fvhoe_rk_combine": (c_int, [c_int, _P, _P, _P, _P, _P, _P, _P, c_longlong, _P]),
    "fvhoe_log_troubles": (c_int, [_P, _P, c_longlong, c_int, c_double, _P]),
    "fvhoe_set_scalar": (c_int, [_P, c_double, _P]),
    "fvhoe_hit_target": (c_int, [_P, _P, _P, c_double, c_int, _P]),
    "fvhoe_advance_time": (c_int, [_P, _P, _P]),
    "fvhoe_log_troubles_gated": (c_int, [_P, _P, c_longlong, c_int, c_double, _P, _P]),
    "fvhoe_advance_time_counted": (c_int, [_P, _P, _P, _P, c_longlong, _P]),
    "fvhoe_log_count": (c_int, [_P, _P, c_longlong, c_int, _P, _P]),
    "fvhoe_log_mean": (c_int, [_P, _P, c_longlong, c_double, _P, _P]),
    "fvhoe_log_magnitudes": (c_int, [_P, _P, _P, _P, c_longlong, _P, _P, _P, _P, c_int, c_longlong, _P, _P]),
    "fvhoe_set_scalar_gated": (c_int, [_P, c_double, _P, _P]),
    "fvhoe_slope_limit": (c_int, [c_int, _P, _P, _P, c_longlong, _P]),
    "fvhoe_stage_rk_static": (c_int, [POINTER(FvhoeStageConfig), _P, _P, _P, _P, POINTER(FvhoeRkFuse), _P]),
    "fvhoe_central_difference": (c_int, [_P, _P, c_longlong, c_longlong, c_longlong, c_double, c_int, _P]),
    "fvhoe_compute_fluxes": (c_int, [_P, _P, _P, c_int, c_longlong, c_int, c_int, _P]),
    "fvhoe_stage_supported": (c_int, [POINTER(FvhoeStageConfig)]),
    "fvhoe_stage_workspace_bytes": (c_int, [POINTER(FvhoeStageConfig), POINTER(c_ulonglong)]),
    "fvhoe_stage": (c_int, [POINTER(FvhoeStageConfig), _P, _P, _P, _P, _P, _P]),
    "fvhoe_stage_rk": (c_int, [POINTER(FvhoeStageConfig), _P, _P, _P, _P, _P, POINTER(FvhoeRkFuse), _P]),
    "fvhoe_stage_timing": (c_int, [c_int]),
    "fvhoe_stage_timing_read": (c_int, [POINTER(c_double), POINTER(c_longlong)]),
    "fvhoe_pred_pack": (c_int, [_P, _P, _P]),
    "fvhoe_pred_pack_planar": (c_int, [_P, _P, _P]),
    "fvhoe_pred_infer": (c_int, [_P, c_int, c_int, _P, c_float, _P, _P, _P]),
    "fvhoe_pred_features": (c_int, [POINTER(FvhoeStageConfig), _P, _P, _P]),
    "fvhoe_pred_conv_forward": (c_int, [_P, _P, _P, _P, c_int, c_int, c_int, c_int, c_int, c_int, _P]),
    "fvhoe_pred_conv_backward": (c_int, [_P, _P, _P, _P, _P, _P, c_int, c_int, c_int, c_int, c_int, _P]),
    "fvhoe_pred_loss": (c_int, [_P, _P, c_float, c_longlong, _P, _P, _P]),
    "fvhoe_pred_adam": (c_int, [_P, _P, _P, _P, c_int, c_float, c_float, c_float, c_float, c_int, _P]),
    "fvhoe_ensemble_workspace_bytes": (c_int, [POINTER(FvhoeStageConfig), c_int, POINTER(c_int), c_int,
                                               POINTER(c_ulonglong)]),
    "fvhoe_ensemble_setup": (c_int, [POINTER(FvhoeStageConfig), c_int, POINTER(c_int), POINTER(c_double), c_int,
                                     _P, _P]),
    "fvhoe_ensemble_step": (c_int, [POINTER(FvhoeStageConfig), c_int, POINTER(c_int), c_int, _P, _P, _P, c_double,
                                    _P, _P]),
    "fvhoe_ensemble_step_persistent": (c_int, [POINTER(FvhoeStageConfig), c_int, POINTER(c_int), c_int, _P, _P, _P,
                                               c_double, _P, _P]),
    "fvhoe_ensemble_status": (c_int, [POINTER(FvhoeStageConfig), c_int, POINTER(c_int), c_int, _P, POINTER(c_int)]),
    "fvhoe_ensemble_step_predicted": (c_int, [POINTER(FvhoeStageConfig), c_int, POINTER(c_int), c_int, _P, _P, _P,
                                              c_double, _P, POINTER(FvhoePredict), _P]),
    "fvhoe_ensemble_pred_logits": (c_int, [POINTER(FvhoeStageConfig), c_int, POINTER(c_int), c_int, _P, _P, _P, _P,
                                           _P, _P]),
    "fvhoe_event_create": (c_int, [POINTER(c_void_p)]),
    "fvhoe_event_destroy": (c_int, [_P]),
    "fvhoe_stage_predicted_workspace_bytes": (c_int, [POINTER(FvhoeStageConfig), POINTER(c_ulonglong)]),
    "fvhoe_stage_predicted": (c_int, [POINTER(FvhoeStageConfig), _P, _P, _P, _P, _P, POINTER(FvhoeRkFuse),
                                      POINTER(FvhoePredict), _P]),
}

# float32 twins of the composed-stage entries (the FP32 bulk form): same arguments
for _name in ("fvhoe_stencil_sweep", "fvhoe_stencil_sweep_zpair", "fvhoe_compute_primitives", "fvhoe_hydro_dt",
              "fvhoe_face_fluxes", "fvhoe_bc_apply", "fvhoe_euler_equation"):
    _SIGNATURES[_name + "_f32"] = _SIGNATURES[_name]
_SIGNATURES["fvhoe_euler_equation_strided"] = (c_int, [_P, POINTER(c_longlong), _P, POINTER(c_longlong), _P,
                                                       POINTER(c_longlong), _P, c_int, c_int, c_int, c_int, c_int,
                                                       c_double, c_double, c_double, _P])
_SIGNATURES["fvhoe_euler_equation_strided_f32"] = _SIGNATURES["fvhoe_euler_equation_strided"]
_SIGNATURES["fvhoe_bc_stencil_sweep"] = (c_int, [_P, _P, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                                 POINTER(c_int), c_int, c_int, c_int, _P])
_SIGNATURES["fvhoe_bc_stencil_sweep_f32"] = _SIGNATURES["fvhoe_bc_stencil_sweep"]
_SIGNATURES["fvhoe_euler_equation_rk"] = (c_int, [_P, POINTER(c_longlong), _P, POINTER(c_longlong), _P,
                                                  POINTER(c_longlong), _P, c_int, c_int, c_int, c_int, c_int,
                                                  c_double, c_double, c_double, POINTER(FvhoeRkFuse), _P])
_SIGNATURES["fvhoe_euler_equation_rk_f32"] = _SIGNATURES["fvhoe_euler_equation_rk"]
_SIGNATURES["fvhoe_stencil_sweep_xypair"] = (c_int, [_P, _P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, _P])
_SIGNATURES["fvhoe_stencil_sweep_xypair_f32"] = _SIGNATURES["fvhoe_stencil_sweep_xypair"]

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m fvhoe_amd.csrc.build` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def check(status: int, what: str):
    if status == OK:
        return
    if status == ERR_ARG:
        raise ValueError(f"{what}: invalid argument")
    if status == ERR_UNSUPPORTED:
        raise NotImplementedError(f"{what}: unsupported configuration")
    if status == ERR_TOO_LARGE:
        raise ValueError(f"{what}: the problem needs more than the library's fixed device scratch of "
                         f"{STATIC_SCRATCH_BYTES} bytes")
    raise RuntimeError(f"{what}: kernel launch failed")


def call(name: str, *args):
    check(getattr(lib(), name)(*args), name)


# ---------------------------------------------------------------- torch glue

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t

        _torch = t
    return _torch


def device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("fvhoe_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return t.device("cuda", t.cuda.current_device())


def stream():
    return c_void_p(torch().cuda.current_stream().cuda_stream)


class OverlapStreams:
    """A side stream (default priority) and two HIP events for running work beside the caller's stream."""

    def __init__(self):
        import weakref

        device()
        self.stream = torch().cuda.Stream()
        fork, join = c_void_p(0), c_void_p(0)
        call("fvhoe_event_create", ctypes.byref(fork))
        call("fvhoe_event_create", ctypes.byref(join))
        self.fork, self.join = fork.value, join.value
        weakref.finalize(self, _destroy_events, lib(), self.fork, self.join)


def _destroy_events(library, *events):
    if _capture_depth > 0:  # destroying an event inside an open capture is not allowed: later
        _deferred.append(lambda: _destroy_events(library, *events))
        return
    for e in events:
        library.fvhoe_event_destroy(c_void_p(e))


# Releases asked for while a graph capture is open, run when the outermost capture closes.  Destroying a
# graph, an event or a stream during a capture aborts the process, whether the owner dies by reference
# count or in a collection.
_capture_depth = 0
_deferred: list = []


def capturing() -> bool:
    return _capture_depth > 0


def defer_release(obj) -> bool:
    """Keep obj alive until the open capture closes (True), or do nothing when no capture is open (False)."""
    if _capture_depth > 0:
        _deferred.append(obj)
        return True
    return False


def _drain_deferred() -> None:
    while _deferred:
        item = _deferred.pop()
        if callable(item) and not isinstance(item, GraphHolder):
            item()
        del item


class GraphHolder:
    """Owner of a captured graph (and whatever else must outlive it): when the holder is released while a
    capture is open, the graph is handed to the deferral list instead of being destroyed on the spot."""

    __slots__ = ("graph", "extra")

    def __init__(self, graph, extra=None):
        self.graph = graph
        self.extra = extra

    def replay(self):
        self.graph.replay()

    def __del__(self):
        g, self.graph = self.graph, None
        if g is not None:
            defer_release(g)


class capture_guard:
    """Context for a HIP graph capture.  Dead objects that hold device resources (solvers in reference
    cycles with their own graphs and tensors) must not be finalised while a capture is open: destroying
    a graph or an event during a capture aborts the process.  On entry: collect, and keep Python's cyclic
    collector off until the capture closes; graphs (GraphHolder) and events (OverlapStreams) released in
    between by reference count are kept on a deferral list and released when the outermost capture
    closes."""

    def __enter__(self):
        import gc

        global _capture_depth
        gc.collect()
        self._was_enabled = gc.isenabled()
        gc.disable()
        _capture_depth += 1
        return self

    def __exit__(self, *exc):
        import gc

        global _capture_depth
        _capture_depth -= 1
        if self._was_enabled:
            gc.enable()
        if _capture_depth == 0:
            _drain_deferred()
        return False


def ptr(t) -> c_void_p:
    return c_void_p(t.data_ptr()) if t is not None else c_void_p(0)


def as_device(x, dtype=None):
    """(device tensor, was_numpy): contiguous float64 on the GPU."""
    t = torch()
    dtype = dtype or t.float64
    if isinstance(x, np.ndarray):
        return t.from_numpy(np.ascontiguousarray(x)).to(device=device(), dtype=dtype), True
    if not isinstance(x, t.Tensor):
        x = t.as_tensor(np.asarray(x))
        return x.to(device=device(), dtype=dtype).contiguous(), True
    if x.device.type != "cuda":
        return x.to(device=device(), dtype=dtype).contiguous(), True
    return x.to(dtype=dtype).contiguous(), False


def as_device_bulk(x):
    """as_device, except that a float32 device tensor (an intermediate of the FP32 bulk form) stays float32."""
    t = torch()
    if isinstance(x, t.Tensor) and x.dtype == t.float32 and x.device.type == "cuda":
        return x.contiguous(), False
    return as_device(x)


def sfx(x) -> str:
    """Entry-name suffix for an array of the bulk type: "_f32" for float32, "" for float64."""
    return "_f32" if x.dtype == torch().float32 else ""


def empty(shape, dtype=None):
    t = torch()
    return t.empty(tuple(int(s) for s in shape), dtype=dtype or t.float64, device=device())


def zeros(shape, dtype=None):
    t = torch()
    return t.zeros(tuple(int(s) for s in shape), dtype=dtype or t.float64, device=device())


def back(x, was_numpy: bool):
    return x.cpu().numpy() if was_numpy else x


def shape4(x):
    s = tuple(x.shape)
    if len(s) != 4:
        raise ValueError(f"expected a 4-D (nvars, nx, ny, nz) array, got shape {s}")
    return s
