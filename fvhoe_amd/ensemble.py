"""Batched ensemble of independent 1-D runs of uneven size on one GPU.

Every instance is an EulerSolver run along x (fvhoe/solver.py:41-1001) with
the shared scheme settings, advanced by SSP-RK3 (fvhoe/ode.py:271-283) with
its own CFL dt, time and _hit_target (fvhoe/ode.py:11-22).  One ensemble
step is a fixed sequence of kernel launches over all instances
(`fvhoe_ensemble_step`); it is captured once in a HIP graph and replayed, and
instances that have reached the target time are skipped inside the kernels,
so the graph never changes.  The FP64 form reproduces each instance's
single-run result bit for bit; bulk_precision="fp32" is the FP32 bulk form.

The tile size (cells per workgroup) is chosen by `autotune`, which times the
captured step for each candidate on the ensemble itself.

With `predictor=` (a TroubledCellPredictor) every step runs the learned
stage: the 2-D network applied to each run seen as a 2-D state constant
along y (which acts as a 1-D network whose kernels are the 2-D kernels summed
over their y taps) pre-flags cells on the step's input; they get MUSCL
fluxes before the candidate in each stage, and the plain check still runs on
every cell (`fvhoe_ensemble_step_predicted`).
"""

from __future__ import annotations

from typing import Sequence

import numpy as np

from fvhoe_amd import _lib
from fvhoe_amd.boundary_conditions import BoundaryCondition
from fvhoe_amd.fv import fv_uniform_meshgen
from fvhoe_amd.slope_limiting import nad_mask, pad_spec

TILES = (64, 128, 256)


class Ensemble:
    """Instances u0[i] of shape (5, n_i) (or (5, n_i, 1, 1)) on the common domain x."""

    def __init__(self, u0: Sequence[np.ndarray], x=(-5.0, 5.0), p: int = 5, gamma: float = 1.4, CFL: float = 0.8,
                 riemann_solver: str = "hllc", bc: BoundaryCondition = None, a_posteriori_slope_limiting=True,
                 slope_limiter: str = "minmod", NAD: float = 1e-2, NAD_mode: str = "global",
                 NAD_range: str = "relative", NAD_vars: tuple = None, PAD: dict = None, SED: bool = True,
                 SED_tolerance: float = 1e-10, csq_floor: float = 1e-10, bulk_precision: str = "fp64",
                 tile: int = 128, persistent: bool = False, ic: Sequence = None, slab_buffer_size: int = 30,
                 predictor=None, count_troubles: bool = False):
        torch = _lib.torch()
        if bulk_precision not in ("fp64", "fp32"):
            raise ValueError(f"bulk_precision must be 'fp64' or 'fp32', got {bulk_precision!r}")
        if riemann_solver not in _lib.RIEMANN:
            raise TypeError(f"Invalid Riemann solver {riemann_solver}")
        arrs = [np.asarray(a, dtype=np.float64).reshape(5, -1) for a in u0]
        self.n = [a.shape[1] for a in arrs]
        self.count = len(arrs)
        self.x_domain = x
        self.h = [(x[1] - x[0]) / n for n in self.n]
        self.bc = BoundaryCondition(x=bc.x, x_value=bc.x_value) if bc is not None else BoundaryCondition(x="outflow")
        c = _lib.FvhoeStageConfig()
        c.n[:] = [1, 1, 1]
        c.p[:] = [p, 0, 0]
        c.h[:] = [1.0, 1.0, 1.0]
        c.gamma, c.cfl, c.csq_floor = float(gamma), float(CFL), float(csq_floor)
        c.bc[:] = self.bc.codes()
        c.riemann = _lib.RIEMANN[riemann_solver]
        c.a_posteriori = int(bool(a_posteriori_slope_limiting))
        c.limiter = 3 if slope_limiter is None else _lib.LIMITER[slope_limiter]
        c.force_trouble = 0
        c.nad_eps = float(NAD)
        c.nad_mode = {"global": 0, "local": 1}[NAD_mode]
        c.nad_range = {"relative": 0, "absolute": 1}[NAD_range]
        c.nad_vars = nad_mask(NAD_vars)
        c.pad_var, c.pad_lo, c.pad_hi = pad_spec({"rho": (0.0, np.inf), "P": (0.0, np.inf)} if PAD is None else PAD)
        c.sed = int(bool(SED))
        c.sed_eps = float(SED_tolerance)
        c.fixed_dt_on = 0
        c.fixed_dt = 0.0
        c.bulk_fp32 = int(bulk_precision == "fp32")
        self._slabs = self._make_slabs(c, p, ic, slab_buffer_size)
        self._cfg = c
        self._n_c = (_lib.ctypes.c_int * self.count)(*self.n)
        self._h_c = (_lib.ctypes.c_double * self.count)(*self.h)
        self.offsets = np.concatenate([[0], np.cumsum(self.n)]).astype(np.int64)
        flat = np.concatenate([a.ravel() for a in arrs])  # instance-major, each (5, n_i)
        self.u = torch.from_numpy(flat).to(_lib.device())
        self.t = _lib.zeros((self.count,))
        self.steps = _lib.zeros((self.count,), dtype=torch.int32)  # steps taken per instance
        self.step_count = 0
        self._graph = None
        self._ws = None
        self.tile = None
        self.persistent = bool(persistent)  # one persistent kernel per RK stage instead of a chain of launches
        # learned stage: the planar network's packed weights; counters [pre-flagged cells, cells the check
        # listed, pre-flagged cells the check did not flag], summed over stages (also without a predictor
        # when count_troubles, for the second-pass share of a learned run)
        self.predictor = predictor
        self._pred = None
        if predictor is not None or count_troubles:
            if predictor is not None and not a_posteriori_slope_limiting:
                raise ValueError("the troubled-cell predictor needs a-posteriori slope limiting")
            if persistent:
                raise NotImplementedError("the learned stage runs as the chain of launches")
            self.predictor_stats = _lib.zeros((3,), dtype=torch.int64)
            pr = _lib.FvhoePredict()
            if predictor is not None:
                self._planar = _lib.empty((_lib.PRED_PLANAR_BYTES,), dtype=torch.uint8)
                _lib.call("fvhoe_pred_pack_planar", _lib.ptr(predictor.theta), _lib.ptr(self._planar), _lib.stream())
                pr.packed = self._planar.data_ptr()
                pr.logit_threshold = float(predictor.logit_threshold)
                pr.infer = 1
            pr.stats = self.predictor_stats.data_ptr()
            self._pred = pr
        self.set_tile(tile)

    # ----------------------------------------------------------------- set-up
    def _make_slabs(self, c, p: int, ic, buffer: int):
        """Ghost values of "dirichlet" / "ic" sides (fvhoe/boundary_conditions.py:18-57, 334-348), per
        instance, as deep as the stage reads: the values at the ghost cell centres of each instance's own
        slab buffer (fvhoe/fv.py:12-86, the mesh EulerSolver builds), an "ic" side from that instance's
        pointwise conservative initial condition ic[i](x, y, z), a "dirichlet" side from its constant or
        callable value at t = 0 (time-dependent values are not supported here)."""
        kinds = self.bc.x
        if not any(k in ("dirichlet", "ic") for k in kinds):
            return None
        if any(k == "dirichlet" and callable(v) for k, v in zip(kinds, self.bc.x_value)):
            raise NotImplementedError("the ensemble takes constant Dirichlet values")
        if "ic" in kinds and (ic is None or len(ic) != self.count):
            raise ValueError("'ic' boundaries need the pointwise initial condition of every instance (ic=[...])")
        depth = _lib.ctypes.c_int(0)
        scheme = _lib.FvhoeStageConfig()
        scheme.p[:] = [p, 0, 0]
        _lib.call("fvhoe_stage_ghost_depth", _lib.ctypes.byref(scheme), _lib.ctypes.byref(depth))
        g = depth.value
        if g > buffer:
            raise ValueError(f"Cannot apply bc to region of thickness {g} with a buffer of thickness {buffer}")
        torch = _lib.torch()
        out = []
        for j, pos in enumerate("lr"):
            if kinds[j] not in ("dirichlet", "ic"):
                out.append(None)
                continue
            vals = np.empty((self.count, 5, g))
            for i in range(self.count):
                _, coords = fv_uniform_meshgen((self.n[i], 1, 1), self.x_domain, (0, 1), (0, 1),
                                               slab_thickness=(buffer, 0, 0))
                values = list(self.bc.x_value)
                if kinds[j] == "ic":
                    values[j] = ic[i]
                bci = BoundaryCondition(x=kinds, x_value=tuple(values), slab_coords=coords)
                vals[i] = bci.slab_values((g, 0, 0), 0, pos, 5, 0.0).reshape(5, g)
            out.append(torch.from_numpy(vals).to(_lib.device()))
        for k, sl in enumerate(out):
            c.slab[k] = sl.data_ptr() if sl is not None else None
        c.slab_g = g
        return out  # the device arrays the kernels read stay alive with the ensemble

    def set_tile(self, tile: int) -> None:
        if tile not in TILES:
            raise ValueError(f"tile must be one of {TILES}")
        nbytes = _lib.ctypes.c_ulonglong(0)
        _lib.call("fvhoe_ensemble_workspace_bytes", _lib.ctypes.byref(self._cfg), self.count, self._n_c, tile,
                  _lib.ctypes.byref(nbytes))
        self._ws = _lib.empty((nbytes.value + 256,), dtype=_lib.torch().uint8)
        _lib.call("fvhoe_ensemble_setup", _lib.ctypes.byref(self._cfg), self.count, self._n_c, self._h_c, tile,
                  self._wsp(), _lib.stream())
        self.tile = tile
        self._graph = None

    def _wsp(self):
        base = self._ws.data_ptr()
        return _lib.ctypes.c_void_p(base + (-base) % 256)

    # ------------------------------------------------------------------- steps
    def _call_step(self, target: float, stream) -> None:
        args = (_lib.ctypes.byref(self._cfg), self.count, self._n_c, self.tile, _lib.ptr(self.u), _lib.ptr(self.t),
                _lib.ptr(self.steps), float(target), self._wsp())
        if self._pred is not None:  # the learned stage (or the plain one with its counters)
            _lib.call("fvhoe_ensemble_step_predicted", *args, _lib.ctypes.byref(self._pred), stream)
        elif self.persistent:  # one persistent kernel per RK stage
            _lib.call("fvhoe_ensemble_step_persistent", *args, stream)
        else:  # the chain of launches per stage
            _lib.call("fvhoe_ensemble_step", *args, stream)

    def step(self, target: float) -> None:
        """One SSP-RK3 step of every instance whose time is below target (no host sync)."""
        self._call_step(target, _lib.stream())
        self.step_count += 1

    def predictor_logits(self):
        """The planar network's logits on the current states (sum n_i floats, instance-major)."""
        if self.predictor is None:
            raise ValueError("no predictor")
        out = _lib.empty((int(self.offsets[-1]),), dtype=_lib.torch().float32)
        _lib.call("fvhoe_ensemble_pred_logits", _lib.ctypes.byref(self._cfg), self.count, self._n_c, self.tile,
                  _lib.ptr(self.u), _lib.ptr(self.t), self._wsp(), _lib.ptr(self._planar), _lib.ptr(out),
                  _lib.stream())
        return out

    def check(self) -> None:
        """Raise if a persistent stage's grid barrier timed out (its results are then not valid)."""
        if not self.persistent:
            return
        err = _lib.ctypes.c_int(0)
        _lib.call("fvhoe_ensemble_status", _lib.ctypes.byref(self._cfg), self.count, self._n_c, self.tile,
                  self._wsp(), _lib.ctypes.byref(err))
        if err.value:
            raise RuntimeError("a grid barrier of the persistent ensemble stage timed out (blocks not all resident)")

    def _capture(self, target: float):
        torch = _lib.torch()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with _lib.capture_guard(), torch.cuda.stream(s):
            # recorded, not run: the state is unchanged by the capture
            with torch.cuda.graph(g, stream=s):
                self._call_step(target, _lib.ctypes.c_void_p(s.cuda_stream))
        torch.cuda.current_stream().wait_stream(s)
        self._graph = (_lib.GraphHolder(g), float(target))
        return self._graph[0]

    def run(self, T: float, graph: bool = True, check_every: int = 64, max_steps: int = 10 ** 7) -> int:
        """Advance every instance to time T; returns the number of ensemble steps taken."""
        torch = _lib.torch()
        steps = 0
        if graph and (self._graph is None or self._graph[1] != float(T)):
            self._capture(T)
        while steps < max_steps:
            for _ in range(check_every):
                if graph:
                    self._graph[0].replay()
                else:
                    self.step(T)
                steps += 1
            tmin = float(self.t.min().item())
            self.check()
            if not np.isfinite(tmin) or tmin >= T:
                break
        torch.cuda.synchronize()
        self.check()
        self.step_count += steps if graph else 0
        return steps

    def autotune(self, T: float, steps: int = 20, tiles=TILES, persistent=(False, True)) -> dict:
        """Time `steps` captured steps for each tile size and each stage form (the chain of launches,
        the persistent kernel) from the current state (restored afterwards) and keep the fastest;
        returns {(tile, "persistent" | "chain"): ms per step}."""
        torch = _lib.torch()
        u0, t0, k0 = self.u.clone(), self.t.clone(), self.steps.clone()
        st0 = self.predictor_stats.clone() if self._pred is not None else None
        if self._pred is not None:
            persistent = (False,)  # the learned stage (and the counting one) runs as the chain
        res = {}
        for pers in persistent:
            self.persistent = bool(pers)
            for tile in tiles:
                self.set_tile(tile)
                g = self._capture(T)
                g.replay()
                torch.cuda.synchronize()
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                for _ in range(steps):
                    g.replay()
                b.record()
                torch.cuda.synchronize()
                self.check()
                res[(tile, "persistent" if pers else "chain")] = a.elapsed_time(b) / steps
                self.u.copy_(u0)
                self.t.copy_(t0)
                self.steps.copy_(k0)
                if st0 is not None:
                    self.predictor_stats.copy_(st0)
        best = min(res, key=res.get)
        self.persistent = best[1] == "persistent"
        self.set_tile(best[0])
        return res

    # ------------------------------------------------------------------ access
    def state(self, i: int) -> np.ndarray:
        """Conservative state (5, n_i) of instance i (host copy)."""
        a, b = self.offsets[i], self.offsets[i + 1]
        return self.u[5 * a:5 * b].reshape(5, b - a).cpu().numpy()

    def times(self) -> np.ndarray:
        return self.t.cpu().numpy()

    def cell_updates(self) -> int:
        """Cells x RK stages performed so far, summed over the instances."""
        k = self.steps.cpu().numpy().astype(np.int64)
        return int(3 * np.dot(k, np.asarray(self.n, dtype=np.int64)))


def shu_osher_ensemble_solver(count: int = 512, seed: int = 0, nmin: int = 256, nmax: int = 8192, p: int = 5,
                              bulk_precision: str = "fp64", tile: int = 128, boundary: str = "ic", **kw) -> Ensemble:
    """The Shu-Osher-type ensemble on [-5, 5]: N ~ U{nmin..nmax}, eps ~ U[0.1, 0.3],
    k ~ U[3, 7] per instance (seeded), p=5 HLLC a-posteriori.  boundary "ic": the ghost cells take the
    initial condition at their centres, as the reference runs Shu-Osher (tests/test_boundary_conditions.py:
    146-224); "outflow": mirror ghosts."""
    from fvhoe_amd import initial_conditions as ics

    insts = ics.shu_osher_ensemble(count, seed=seed, nmin=nmin, nmax=nmax)
    u0 = [ics.shu_osher_instance(n, e, k) for n, e, k in insts]
    if boundary == "ic":
        point = [ics.ShuOsherInstance(eps=e, k=k).build_ic() for _, e, k in insts]
        kw["bc"] = BoundaryCondition(x="ic")
        kw["ic"] = [(lambda f: (lambda x, y, z: ics.primitives_to_conservatives(f(x, y, z), 1.4)))(f) for f in point]
    elif boundary != "outflow":
        raise ValueError(f"boundary must be 'ic' or 'outflow', got {boundary!r}")
    return Ensemble(u0, x=(-5.0, 5.0), p=p, bulk_precision=bulk_precision, tile=tile, **kw)
