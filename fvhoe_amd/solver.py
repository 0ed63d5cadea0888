"""EulerSolver on MI355X (fvhoe/solver.py:41-1001).

Same constructor arguments, steppers and snapshot format as the reference.
Each RK stage runs either

* the fused stage (`fvhoe_stage` in the C-ABI: a handful of LDS-tiled HIP
  kernels) for 1-D runs along x and 2-D runs in x-y with 5 variables, the same
  degree on both axes and the simple face broadcast, or
* the composed stage, which follows hydrofluxes / revise_fluxes /
  euler_equation (fvhoe/solver.py:456-751) call for call, one HIP kernel per
  array function, for every other configuration (3-D, passives, mixed
  degrees, convex blending, x-z / y-z planes).

Both produce the reference's FP64 results bit for bit.
"""

from __future__ import annotations

import json
import os
import pickle
import shutil
from typing import Callable, Dict, Iterable, Tuple

import numpy as np

from fvhoe_amd import _lib
from fvhoe_amd.boundary_conditions import BoundaryCondition
from fvhoe_amd.fv import (conservative_interpolation, fv_uniform_meshgen, interpolate_cell_centers,
                          interpolate_fv_averages, sweep_chain)
from fvhoe_amd.fv import transverse_reconstruction
from fvhoe_amd.hydro import HydroState, compute_primitives, hydro_dt_device
from fvhoe_amd.initial_conditions import (InitialCondition, Square, device_mesh, fv_average,
                                          primitives_to_conservatives, uniform_mesh)
from fvhoe_amd.ode import ODE
from fvhoe_amd.slope_limiting import (MUSCL_interpolations, broadcast_to_troubled_interfaces, detect_troubled_cells,
                                      nad_mask, pad_spec)
from fvhoe_amd.stencils import get_stencil_size


MAG_SLOTS = 4  # stages whose violation magnitudes the fused stage keeps in slots (RK4's four)


def _jsonable(v) -> bool:
    if v is None or isinstance(v, (bool, int, float, str)):
        return not isinstance(v, float) or np.isfinite(v)
    if isinstance(v, (tuple, list)):
        return all(_jsonable(x) for x in v)
    if isinstance(v, dict):
        return all(isinstance(k, str) and _jsonable(x) for k, x in v.items())
    return False


def timing_report(timer: dict) -> str:
    """Category / seconds table of the run's timer (the layout of fvhoe/timer.py:78-110)."""
    cats = list(timer)
    times = [f"{float(timer[c]):.2f}" for c in cats]
    wc = max([len(c) for c in cats] or [len("Category")])
    wt = max([len(t) for t in times] + [len("Time (s)")])
    out = f"{'Category':<{wc}} {'Time (s)':<{wt}}\n" + "-" * (wc + wt + 1) + "\n"
    return out + "".join(f"{c:<{wc}} {t:>{wt}}\n" for c, t in zip(cats, times))


def commit_details() -> dict:
    """The package's git commit (hash, author, date, branch), or the error git reported."""
    import subprocess

    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    try:
        r = subprocess.run(["git", "-C", repo, "log", "-1", "--pretty=format:%H|%an|%ai|%D"],
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, check=True)
    except (OSError, subprocess.CalledProcessError) as e:
        return {"error": f"An error occurred: {getattr(e, 'stderr', str(e)) or e}".strip()}
    parts = r.stdout.strip().split("|")
    parts += [""] * (4 - len(parts))
    branch = parts[3].split(",")[0].strip().split()[-1] if parts[3].strip() else None
    return {"commit_hash": parts[0], "author_name": parts[1], "commit_date": parts[2], "branch_name": branch}


class _TroubleLog(dict):
    """The solver's log arrays.  Trouble flags are counted in integers between resets and turned into
    the mean at the step's reset; reading "mean trouble" before a reset writes the running sum the
    reference holds at that point (counts are exact integer sums, so the doubles are equal)."""

    def __init__(self, solver, arrays):
        super().__init__(arrays)
        self._solver = solver

    def __getitem__(self, key):
        out = dict.__getitem__(self, key)
        s = self._solver
        if key == "mean trouble" and s._count_pending:
            _lib.call("fvhoe_log_mean", _lib.ptr(out), _lib.ptr(s._trouble_count), out.numel(), 1.0, None,
                      _lib.stream())
        elif key in ("mean NAD mag", "mean PAD mag") and s._mag is not None:
            s._mag_materialize()
        return out


class EulerSolver(ODE):
    def __init__(
        self,
        w0: InitialCondition = Square(),
        w0_passives: Dict[str, Callable] = None,
        nx: int = 1, ny: int = 1, nz: int = 1,
        px: int = 0, py: int = 0, pz: int = 0,
        x: Tuple[float, float] = (0, 1), y: Tuple[float, float] = (0, 1), z: Tuple[float, float] = (0, 1),
        CFL: float = 0.8,
        fixed_dt: float = None,
        gamma: float = 5 / 3,
        bc: BoundaryCondition = None,
        riemann_solver: str = "hllc",
        conservative_ic: bool = False,
        fv_ic: bool = False,
        fixed_primitive_variables: Iterable = None,
        a_posteriori_slope_limiting: bool = False,
        slope_limiter: str = "minmod",
        force_trouble: bool = False,
        NAD: float = 1e-2,
        NAD_mode: str = "global",
        NAD_range: str = "relative",
        NAD_vars: tuple = None,
        PAD: dict = None,
        SED: bool = True,
        SED_tolerance: float = 1e-10,
        convex: bool = False,
        density_floor: float = None,
        pressure_floor: float = None,
        csq_floor: float = 1e-10,
        progress_bar: bool = True,
        snapshots_as_fv_averages: bool = True,
        snapshot_helper_function: callable = None,
        slab_buffer_size: int = 30,
        cupy: bool = False,
        fused: bool = None,
        bulk_precision: str = "fp64",
        predictor=None,
        predictor_every_stage: bool = False,
        predictor_overlap: bool = False,
        ic_device: bool = False,
    ):
        if w0 is None:
            w0 = Square()
        if riemann_solver not in _lib.RIEMANN:
            raise TypeError(f"Invalid Riemann solver {riemann_solver}")
        if slope_limiter not in _lib.LIMITER:
            raise ValueError(f"Unknown slope limiter: {slope_limiter}")
        if bulk_precision not in ("fp64", "fp32"):
            raise ValueError(f"bulk_precision must be 'fp64' or 'fp32', got {bulk_precision!r}")
        # "fp32": reconstruction sweeps and Riemann fluxes in FP32 (fused stage only); pressure recovery,
        # trouble tests, divergence and the RK accumulation stay FP64
        self.bulk_precision = bulk_precision
        self.csq_floor = csq_floor
        self.density_floor = density_floor  # stored, never applied (as in the reference)
        self.pressure_floor = pressure_floor
        self.gamma = gamma
        self.riemann_solver_name = riemann_solver
        self.snapshot_helper_function = snapshot_helper_function
        self.snapshots_as_fv_averages = snapshots_as_fv_averages
        self.CFL, self.fixed_dt = CFL, fixed_dt
        self.x_domain, self.y_domain, self.z_domain = x, y, z
        self.n = (nx, ny, nz)
        self.p = (px, py, pz)
        self.h = ((x[1] - x[0]) / nx, (y[1] - y[0]) / ny, (z[1] - z[0]) / nz)
        self.X, self.Y, self.Z = uniform_mesh(self.n, x, y, z)
        self.x, self.y, self.z = self.X[:, 0, 0], self.Y[0, :, 0], self.Z[0, 0, :]
        self.xdim = not (nx == 1 and px == 0)
        self.ydim = not (ny == 1 and py == 0)
        self.zdim = not (nz == 1 and pz == 0)
        self.active = (self.xdim, self.ydim, self.zdim)
        self.dims = "".join(d for d, a in zip("xyz", self.active) if a)
        self.ndim = len(self.dims)

        passives = tuple(w0_passives.keys()) if w0_passives is not None else ()
        self.hydro_state = HydroState(passive_scalars=passives)
        self.nvars = self.hydro_state.nvars
        ic = w0.build_ic(passives, w0_passives)
        # the cell averages are built on the GPU with ic_device (IC functions written for NumPy or torch,
        # as every class of fvhoe_amd.initial_conditions is); the mesh holds the same doubles either way
        Xi, Yi, Zi = device_mesh(self.n, x, y, z) if ic_device else (self.X, self.Y, self.Z)
        # self.u0: pointwise conservative IC f(x, y, z) where the IC defines one (fvhoe/solver.py:278-299)
        self.conservative_ic, self.fv_ic = conservative_ic, fv_ic
        # self.w0: the pointwise primitive IC where the IC gives primitives (fvhoe/solver.py:281-299)
        self.w0 = ic if not conservative_ic else NotImplemented
        if conservative_ic and fv_ic:
            self.u0 = NotImplemented
            u0 = ic(Xi, Yi, Zi)
        elif fv_ic:
            raise ValueError("Can't have finite volume averages without conservative IC.")
        else:
            point = ic if conservative_ic else (lambda x, y, z: primitives_to_conservatives(ic(x, y, z), gamma))
            self.u0 = point
            u0 = fv_average(point, Xi, Yi, Zi, self.h, self.p)
        super().__init__(u0, progress_bar=progress_bar, cupy=cupy)
        self.f_evaluation_count = 0

        # boundaries (fvhoe/solver.py:314-362): slab kinds read their ghost values at the centres of a
        # slab_buffer_size-thick ghost buffer; "ic" sides take the pointwise initial condition
        self.slab_buffer_size = slab_buffer_size
        if bc is None:
            self.bc = BoundaryCondition()
        else:
            slab_coords = None
            if bc.has_slabs():
                sizes = tuple(slab_buffer_size if a else 0 for a in self.active)
                _, slab_coords = fv_uniform_meshgen(self.n, x, y, z, slab_thickness=sizes)
            self.bc = BoundaryCondition(x=bc.x, y=bc.y, z=bc.z, x_value=bc.x_value, y_value=bc.y_value,
                                        z_value=bc.z_value, slab_coords=slab_coords)
            for d in "xyz":
                for j, pos in enumerate("lr"):
                    if getattr(self.bc, d)[j] == "ic":
                        self.bc.reset_value(d, pos, self.u0)
        self.a_posteriori_slope_limiting = a_posteriori_slope_limiting
        self.slope_limiter = slope_limiter
        self.force_trouble = force_trouble
        self.NAD, self.NAD_mode, self.NAD_range = NAD, NAD_mode, NAD_range
        self.NAD_vars = NAD_vars if isinstance(NAD_vars, tuple) or NAD_vars is None else tuple(NAD_vars)
        self.PAD = {"rho": (0.0, np.inf), "P": (0.0, np.inf)} if PAD is None else PAD
        self.SED, self.SED_tolerance, self.convex = SED, SED_tolerance, convex
        self.trouble_counter = 1  # fvhoe/solver.py:400
        nx_, ny_, nz_ = self.n
        self.log = _TroubleLog(self, {k: _lib.zeros((nx_, ny_, nz_)) for k in ("mean trouble", "mean NAD mag",
                                                                               "mean PAD mag")})
        self._zero_like_log = _lib.zeros((nx_, ny_, nz_))
        self._log_written = {}  # log arrays that have received magnitudes from the composed stage
        self._trouble_count = _lib.zeros((nx_, ny_, nz_), dtype=_lib.torch().int32)  # flags counted since a reset
        self._count_pending = False
        self._one = _lib.zeros((1,))
        self._one.fill_(1.0)

        self._cfg = self._stage_config()
        supported = bool(_lib.lib().fvhoe_stage_supported(self._cfg)) and self.nvars == 5 and not convex
        self.fused = supported if fused is None else (bool(fused) and supported)
        if fused and not supported:
            raise NotImplementedError("the fused stage does not cover this configuration")
        # FP32 bulk form: the fused stage (1-D x or 2-D x-y), or the composed stage with float32 intermediates
        # (5 variables, no a-posteriori limiting, no slab boundaries: the 3-D form)
        self._bulk32 = bulk_precision == "fp32" and not self.fused
        if self._bulk32 and (a_posteriori_slope_limiting or self.nvars != 5 or self.bc.has_slabs()):
            raise NotImplementedError("the FP32 bulk form runs on the fused stage (1-D x or 2-D x-y, 5 variables) or, "
                                      "without a-posteriori limiting or slab boundaries, on the composed stage")
        self._ws = None
        self._trouble = _lib.empty(self.n, dtype=_lib.torch().int8) if a_posteriori_slope_limiting else None
        # violation magnitudes of the fused stage: the check appends (cell, NAD magnitude, PAD magnitude) of
        # its flagged cells to one slot per stage; reading the log turns the slots of the current
        # accumulation into the reference's dense running sums (or means, after a reset)
        self._mag = None
        self._mag_used, self._mag_closed, self._mag_base_dense, self._mag_acc = 0, False, False, None
        if self.fused and a_posteriori_slope_limiting:
            torch_, nc = _lib.torch(), int(np.prod(self.n))
            self._mag = dict(count=_lib.zeros((MAG_SLOTS,), dtype=torch_.int32),
                             idx=_lib.empty((MAG_SLOTS, nc), dtype=torch_.int32), nad=_lib.empty((MAG_SLOTS, nc)),
                             pad=_lib.empty((MAG_SLOTS, nc)), div=_lib.zeros((1,)))
            c = self._cfg
            c.mag_count, c.mag_idx = self._mag["count"].data_ptr(), self._mag["idx"].data_ptr()
            c.mag_nad, c.mag_pad = self._mag["nad"].data_ptr(), self._mag["pad"].data_ptr()
        # learned troubled-cell predictor (fvhoe_amd.predictor): pre-flagged cells get MUSCL fluxes before
        # the candidate is formed; the plain check still judges every cell
        self.predictor = predictor
        self.predictor_every_stage = bool(predictor_every_stage)
        # the network of a step's first stage runs on a side stream beside that stage's other kernels;
        # its flags apply from the step's second stage on
        self.predictor_overlap = bool(predictor_overlap) and not self.predictor_every_stage
        self.predictor_stats = None
        self._overlap = None
        if predictor is not None:
            if not (self.fused and self.ndim == 2 and self.xdim and self.ydim and a_posteriori_slope_limiting):
                raise NotImplementedError("the predictor runs in the fused 2-D a-posteriori stage")
            if self.bc.has_slabs():
                raise NotImplementedError("the predictor's features take periodic / mirror / free boundaries only")
            self.predictor_stats = _lib.zeros((3,), dtype=_lib.torch().int64)
            if self.predictor_overlap:
                self._overlap = _lib.OverlapStreams()

    # ------------------------------------------------------------------ config
    def _stage_config(self):
        c = _lib.FvhoeStageConfig()
        c.n[:] = list(self.n)
        c.p[:] = list(self.p)
        c.h[:] = list(self.h)
        c.gamma, c.cfl, c.csq_floor = float(self.gamma), float(self.CFL), float(self.csq_floor)
        c.bc[:] = self.bc.codes()
        c.riemann = _lib.RIEMANN[self.riemann_solver_name]
        c.a_posteriori = int(bool(self.a_posteriori_slope_limiting))
        c.limiter = 3 if self.slope_limiter is None else _lib.LIMITER[self.slope_limiter]
        c.force_trouble = int(bool(self.force_trouble))
        c.nad_eps = float(self.NAD)
        if self.NAD_mode not in ("global", "local"):
            raise ValueError(f"Unknown mode: {self.NAD_mode}")
        if self.NAD_range not in ("relative", "absolute"):
            raise ValueError(f"Unknown range_type: {self.NAD_range}")
        c.nad_mode = 0 if self.NAD_mode == "global" else 1
        c.nad_range = 0 if self.NAD_range == "relative" else 1
        c.nad_vars = nad_mask(self.NAD_vars)
        c.pad_var, c.pad_lo, c.pad_hi = pad_spec(self.PAD)
        c.sed = int(bool(self.SED))
        c.sed_eps = float(self.SED_tolerance)
        c.fixed_dt_on = int(self.fixed_dt is not None)
        c.fixed_dt = float(self.fixed_dt or 0.0)
        c.bulk_fp32 = int(self.bulk_precision == "fp32")
        self._set_stage_slabs(c)
        return c

    def _set_stage_slabs(self, c) -> None:
        """Ghost slabs of the fused stage's Dirichlet / "ic" sides, as deep as the stage reads, evaluated at
        the solver's time (the reference fills every stage's ghosts at self.t, fvhoe/solver.py:784)."""
        if not self.bc.has_slabs() or self.zdim or not (1 <= self.p[0] <= 8):
            return
        depth = _lib.ctypes.c_int(0)
        _lib.call("fvhoe_stage_ghost_depth", _lib.ctypes.byref(c), _lib.ctypes.byref(depth))
        gw = (depth.value if self.xdim else 0, depth.value if self.ydim else 0, 0)
        slabs = self.bc.device_slabs(gw, 5, self.t)
        self._slab_refs = slabs  # the kernels (and captured graphs) read these tensors
        self._slab_version = self.bc.version
        for k in range(6):
            c.slab[k] = slabs[k].data_ptr() if slabs[k] is not None else None
        c.slab_g = depth.value
        c.slab_pad[:] = [gw[0], gw[1], 0]
        self._slab_t = self.t

    def _needs_host_clock(self) -> bool:
        return self.bc.time_dependent()

    def _before_step(self) -> None:
        """Refresh time-dependent boundary slabs of the fused stage to the step's start time."""
        if self.fused and self.bc.time_dependent() and self._slab_t != self.t:
            self._set_stage_slabs(self._cfg)

    def _fused_stage_prologue(self) -> None:
        """Host bookkeeping before a fused stage: boundary values changed by bc.reset_value are fetched
        again (captured steps read the old slabs, so they are dropped), and the stage's magnitude slot is
        chosen."""
        if self.bc.has_slabs() and getattr(self, "_slab_version", self.bc.version) != self.bc.version:
            self._set_stage_slabs(self._cfg)
            self._graphs = {}
        if self._mag is None:
            return
        if self._mag_closed:  # the first stage after a reset starts a new sum (the reference assigns)
            self._mag_used, self._mag_closed, self._mag_base_dense = 0, False, False
        if self._mag_used == MAG_SLOTS:  # a long run of stages without a reset: fold the slots first
            self._mag_fold()
        self._cfg.mag_slot = self._mag_used
        self._mag_used += 1

    def _mag_call(self, nad_out, pad_out, divisor):
        m = self._mag
        base = self._mag_acc if self._mag_base_dense else (None, None)
        nc = m["idx"].shape[1]
        _lib.call("fvhoe_log_magnitudes", _lib.ptr(nad_out), _lib.ptr(pad_out), _lib.ptr(base[0]), _lib.ptr(base[1]),
                  nc, _lib.ptr(m["count"]), _lib.ptr(m["idx"]), _lib.ptr(m["nad"]), _lib.ptr(m["pad"]),
                  self._mag_used, nc, _lib.ptr(divisor), _lib.stream())

    def _mag_fold(self) -> None:
        """Partial sums of the filled slots into the dense accumulator; the slots are free again."""
        if self._mag_acc is None:
            self._mag_acc = (_lib.empty(self.n), _lib.empty(self.n))
        self._mag_call(self._mag_acc[0], self._mag_acc[1], None)
        self._mag_base_dense, self._mag_used = True, 0

    def _mag_materialize(self) -> None:
        """The log's "mean NAD mag" / "mean PAD mag" from the slots: the running sums since the last
        reset, or their means when a reset closed them (the divisor is the device scalar the reset set)."""
        if self._mag_used == 0 and not self._mag_base_dense and not self._mag_closed:
            return  # no stage since construction: the log holds its initial zeros
        lg = self.log
        self._mag_call(dict.__getitem__(lg, "mean NAD mag"), dict.__getitem__(lg, "mean PAD mag"),
                       self._mag["div"] if self._mag_closed else None)

    # ------------------------------------------------------------------- stage
    def stage(self, u, k_out, dt_out) -> None:
        """One RHS evaluation into k_out, this stage's dt into dt_out[0] (device)."""
        self.f_evaluation_count += 1
        if self.fused:
            self._fused_stage_prologue()
        if self.predictor is not None:
            _lib.call("fvhoe_stage_predicted", _lib.ctypes.byref(self._cfg), _lib.ptr(u), _lib.ptr(k_out),
                      _lib.ptr(dt_out), _lib.ptr(self._trouble), self._workspace(), None,
                      _lib.ctypes.byref(self._predict_args(True)), _lib.stream())
            self._log_troubles_device(self._trouble)
        elif self.fused:
            _lib.call("fvhoe_stage", _lib.ctypes.byref(self._cfg), _lib.ptr(u), _lib.ptr(k_out), _lib.ptr(dt_out),
                      _lib.ptr(self._trouble), self._workspace(), _lib.stream())
            if self._trouble is not None:
                self._log_troubles_device(self._trouble)
        else:
            dt, (F, G, H) = self.hydrofluxes(u, self.p, views=not self.a_posteriori_slope_limiting, dt_out=dt_out)
            if self.a_posteriori_slope_limiting:
                F, G, H = self.revise_fluxes(u, F, G, H, dt, self.force_trouble)
            self._euler_equation_into(F, G, H, k_out)

    def _fused_rk(self) -> bool:
        # the composed stage without a-posteriori limiting ends in the divergence + RK combination kernel
        return self.fused or not self.a_posteriori_slope_limiting

    def _chunkable(self) -> bool:
        """Steps back to back on the device: the fused stage only (the composed stage takes no device
        target time)."""
        return self.fused and not self._needs_host_clock()

    def _composed_stage_rk(self, u, k_out, dt_out, kind, out, u0, k0, k1, k2, dt_step, set_dt_step, t_dev, target):
        """The composed stage (no a-posteriori limiting) with the divergence and the RK combination in one
        kernel; the step's dt is clipped to the target first, as the separate combine kernels take it."""
        _, (F, G, H) = self.hydrofluxes(u, self.p, views=True, dt_out=dt_out)
        if set_dt_step:
            _lib.call("fvhoe_hit_target", _lib.ptr(dt_step), _lib.ptr(dt_out), _lib.ptr(t_dev),
                      float(target) if target is not None else 0.0, int(target is not None), _lib.stream())
        rk = _lib.FvhoeRkFuse()
        rk.kind = kind
        rk.out, rk.u0 = out.data_ptr(), u0.data_ptr()
        rk.k0 = k0.data_ptr() if k0 is not None else None
        rk.k1 = k1.data_ptr() if k1 is not None else None
        rk.k2 = k2.data_ptr() if k2 is not None else None
        rk.dt_step = dt_step.data_ptr()
        nv, n0, n1, n2 = u.shape
        active = sum(1 << d for d in range(3) if self.active[d])
        sfx, args, _keep = self._flux_args(F, G, H, u.shape)
        _lib.call("fvhoe_euler_equation_rk" + sfx, *args, _lib.ptr(k_out), nv, n0, n1, n2, active, -(1 / self.h[0]),
                  -(1 / self.h[1]), -(1 / self.h[2]), _lib.ctypes.byref(rk), _lib.stream())
        return True

    def stage_rk(self, u, k_out, dt_out, kind, out, u0, k0=None, k1=None, k2=None, dt_step=None,
                 set_dt_step=False, t_dev=None, target=None, target_dev=None) -> bool:
        """A fused stage whose last kernel also writes the RK combination `kind`
        into `out` (fvhoe/ode.py:245-298), so the state is touched once per stage."""
        self.f_evaluation_count += 1
        if not self.fused:
            if target_dev is not None:
                raise NotImplementedError("the composed stage takes a host target time")
            return self._composed_stage_rk(u, k_out, dt_out, kind, out, u0, k0, k1, k2, dt_step, set_dt_step,
                                           t_dev, target)
        self._fused_stage_prologue()
        rk = _lib.FvhoeRkFuse()
        rk.kind = kind
        rk.out, rk.u0 = out.data_ptr(), u0.data_ptr()
        rk.k0 = k0.data_ptr() if k0 is not None else None
        rk.k1 = k1.data_ptr() if k1 is not None else None
        rk.k2 = k2.data_ptr() if k2 is not None else None
        rk.dt_step = dt_step.data_ptr()
        rk.set_dt_step = int(bool(set_dt_step))
        rk.t = t_dev.data_ptr() if t_dev is not None else None
        rk.target = float(target) if target is not None else 0.0
        rk.has_target = int(target is not None)
        rk.target_ptr = target_dev.data_ptr() if target_dev is not None else None
        if self.predictor is not None:
            pa = self._predict_args(bool(set_dt_step) or self.predictor_every_stage, overlap=self._overlap)
            _lib.call("fvhoe_stage_predicted", _lib.ctypes.byref(self._cfg), _lib.ptr(u), _lib.ptr(k_out),
                      _lib.ptr(dt_out), _lib.ptr(self._trouble), self._workspace(), _lib.ctypes.byref(rk),
                      _lib.ctypes.byref(pa), _lib.stream())
        else:
            _lib.call("fvhoe_stage_rk", _lib.ctypes.byref(self._cfg), _lib.ptr(u), _lib.ptr(k_out),
                      _lib.ptr(dt_out), _lib.ptr(self._trouble), self._workspace(), _lib.ctypes.byref(rk),
                      _lib.stream())
        if self._trouble is not None:
            self._log_troubles_device(self._trouble, gate=dt_step)
        return True

    def _predict_args(self, infer: bool, overlap=None):
        pa = _lib.FvhoePredict()
        pa.packed = self.predictor.packed.data_ptr()
        pa.logit_threshold = float(self.predictor.logit_threshold)
        pa.infer = int(bool(infer))
        pa.stats = self.predictor_stats.data_ptr()
        if overlap is not None:
            pa.overlap = 1
            pa.side_stream = overlap.stream.cuda_stream
            pa.ev_fork, pa.ev_join = overlap.fork, overlap.join
        return pa

    def predictor_features(self):
        """Predictor features (4, nx+12, ny+12) float32 of the input of the last fused stage."""
        X = _lib.empty((4, self.n[0] + 12, self.n[1] + 12), dtype=_lib.torch().float32)
        _lib.call("fvhoe_pred_features", _lib.ctypes.byref(self._cfg), self._workspace(), _lib.ptr(X), _lib.stream())
        return X

    def _workspace(self):
        if self._ws is None:
            nbytes = _lib.ctypes.c_ulonglong(0)
            fn = "fvhoe_stage_predicted_workspace_bytes" if self.predictor is not None else "fvhoe_stage_workspace_bytes"
            _lib.call(fn, _lib.ctypes.byref(self._cfg), _lib.ctypes.byref(nbytes))
            self._ws = _lib.empty((nbytes.value + 256,), dtype=_lib.torch().uint8)
        base = self._ws.data_ptr()
        return _lib.ctypes.c_void_p(base + (-base) % 256)

    def hydrodynamics(self, u):
        """(dt, dudt) of one stage (fvhoe/solver.py:429-454)."""
        x, was_np = _lib.as_device(u)
        k = _lib.empty(x.shape)
        dt = _lib.empty((1,))
        self.stage(x, k, dt)
        return float(dt.item()), _lib.back(k, was_np)

    # ------------------------------------------------------- composed stage
    def interpolate_primitives_from_conservatives(self, u, p, gw=(0, 0, 0), fv_average: bool = True):
        """fvhoe/solver.py:753-792."""
        cost = [(2 if fv_average else 1) * get_stencil_size(pi) for pi in p]
        g = tuple(int(gw[i]) + cost[i] for i in range(3))
        dtype = _lib.torch().float32 if self._bulk32 else None
        # the ghost fill and the x c-sweep in one pass where the fused kernel applies (large grids)
        first = self.bc.apply_swept(u, g, "c", p[0], 1, dtype=dtype) if p[0] > 0 else None
        if first is not None:
            centres = sweep_chain(first, "c", ((p[1], 2), (p[2], 3)))
        else:
            centres = interpolate_cell_centers(self.bc.apply(u, g, t=self.t, dtype=dtype), p)
        w = compute_primitives(self.hydro_state, centres, self.gamma)
        return interpolate_fv_averages(w, p) if fv_average else w

    def hydrofluxes(self, u, p, slope_limiter: str = None, timer_prefix: str = "", views: bool = False,
                    dt_out=None):
        """dt (device), (F, G, H) (fvhoe/solver.py:456-651).  views=True returns the trimmed fluxes as views
        into the untrimmed arrays (for the strided divergence), else contiguous copies; dt_out (1-element
        device tensor) receives dt when given."""
        if slope_limiter is None:
            limit = False
        elif max(p) <= 1:
            limit = True
        else:
            raise ValueError("Slope limiting is only implemented for degree 1 polynomial interpolation.")
        pmax = max(p)
        cm = get_stencil_size(pmax)
        gw = max(cm + 1, 2 * cm)
        w_bc = self.interpolate_primitives_from_conservatives(u, p, tuple(gw if a else 0 for a in self.active))
        if self.fixed_dt is None:
            dt = hydro_dt_device(w_bc, min(self.h), self.ndim, self.CFL, self.gamma, self.csq_floor, out=dt_out)
        else:
            dt = _lib.empty((1,)) if dt_out is None else dt_out
            dt.fill_(float(self.fixed_dt))
        order = {0: (3, 2), 1: (1, 3), 2: (2, 1)}
        tr_order = {0: (2, 3), 1: (1, 3), 2: (1, 2)}
        out = []
        rs = _lib.RIEMANN[self.riemann_solver_name]
        for d in range(3):
            if not self.active[d]:
                out.append(None)
                continue
            a = d + 1
            wt = sweep_chain(w_bc, "c", tuple((p[ax - 1], ax) for ax in order[d]))
            if limit and p[d] == 1:
                wl_face, wr_face = MUSCL_interpolations(wt, a, slope_limiter)
                f = self._riemann_faces(wr_face, wl_face, a, rs)
            elif wt.shape[0] == 5 and wt.numel() < (1 << 31):
                # face states and Riemann solve in one pass (the same operations as the three kernels)
                shape = list(wt.shape)
                shape[a] -= 2 * get_stencil_size(p[d]) + 1
                f = _lib.empty(shape, dtype=wt.dtype)
                _lib.call("fvhoe_face_fluxes" + _lib.sfx(wt), _lib.ptr(wt), _lib.ptr(f), *wt.shape[1:], a, p[d], rs,
                          float(self.gamma), float(self.csq_floor), _lib.stream())
            elif self._bulk32:
                raise NotImplementedError("the FP32 composed stage needs 5 variables and < 2^31 elements per array")
            else:
                wl_face = conservative_interpolation(wt, p[d], a, "l")
                wr_face = conservative_interpolation(wt, p[d], a, "r")
                f = self._riemann_faces(wr_face, wl_face, a, rs)
            f = sweep_chain(f, "t", tuple((p[ax - 1], ax) for ax in tr_order[d]))
            excess = [0, 0, 0, 0]
            for e in range(3):
                if self.active[e]:
                    excess[e + 1] = gw - get_stencil_size(p[e]) - 1 if e == d else gw - 2 * get_stencil_size(p[e])
            f = f[tuple(slice(x or None, -x or None) for x in excess)]
            out.append(f if views else f.contiguous())
        return dt, tuple(out)

    def _riemann_faces(self, wr_face, wl_face, axis, rs):
        """RS(wr_face[:-1], wl_face[1:]) along axis (fvhoe/solver.py:573-606)."""
        nv, n0, n1, n2 = wr_face.shape
        shape = [nv, n0, n1, n2]
        shape[axis] -= 1
        out = _lib.empty(shape)
        if nv == 5:
            _lib.call("fvhoe_riemann_faces", _lib.ptr(wr_face), _lib.ptr(wl_face), _lib.ptr(out), n0, n1, n2, axis, rs,
                      float(self.gamma), float(self.csq_floor), _lib.stream())
        else:
            sl_l = [slice(None)] * 4
            sl_r = [slice(None)] * 4
            sl_l[axis] = slice(None, -1)
            sl_r[axis] = slice(1, None)
            a = wr_face[tuple(sl_l)].contiguous()
            b = wl_face[tuple(sl_r)].contiguous()
            nc = a[0].numel()
            _lib.call("fvhoe_riemann", _lib.ptr(a), _lib.ptr(b), _lib.ptr(out), nv, nc, rs, axis - 1,
                      float(self.gamma), float(self.csq_floor), _lib.stream())
        return out

    def revise_fluxes(self, u, F, G, H, dt, force_trouble: bool = False):
        """A-posteriori check and MUSCL fallback blend (fvhoe/solver.py:653-732);
        dt is the stage's device dt.  Returns the revised (F, G, H)."""
        L = self.euler_equation(F, G, H)
        ustar = _lib.empty(u.shape)
        _lib.call("fvhoe_rk_combine", _lib.RK_AXPY, _lib.ptr(ustar), _lib.ptr(u), _lib.ptr(L), None, None, None,
                  _lib.ptr(dt), u.numel(), _lib.stream())
        gw = tuple(3 if a else 0 for a in self.active)
        w = self.interpolate_primitives_from_conservatives(u, self.p, gw)
        ws = self.interpolate_primitives_from_conservatives(ustar, self.p, gw)
        trouble, nad_mag, pad_mag = detect_troubled_cells(
            w[:5].contiguous(), ws[:5].contiguous(), self.dims, self.NAD, self.NAD_mode, self.NAD_range,
            self.NAD_vars, self.PAD, self.SED, self.SED_tolerance)
        if force_trouble:
            trouble = _lib.torch().ones_like(trouble)
        self.log_troubles(trouble, nad_mag, pad_mag)
        if bool(trouble.any().item()):
            p1 = tuple(int(pi > 0) for pi in self.p)
            _, (Fl, Gl, Hl) = self.hydrofluxes(u, p1, slope_limiter=self.slope_limiter)
            per = self.bc.periodic()
            tx, ty, tz = broadcast_to_troubled_interfaces(trouble, self.dims, self.convex, *per)
            F = self._blend(F, Fl, tx)
            G = self._blend(G, Gl, ty)
            H = self._blend(H, Hl, tz)
        return F, G, H

    def _blend(self, F, Fl, th):
        if F is None:
            return None
        out = F.clone()
        th = th.contiguous()
        _lib.call("fvhoe_blend", _lib.ptr(out), _lib.ptr(Fl), _lib.ptr(th), F.shape[0], th.numel(), _lib.stream())
        return out

    def _flux_args(self, F, G, H, shape):
        """(pointer, element strides) per face-flux array for the strided divergence entries."""
        nv, n0, n1, n2 = shape
        first = next(f for f in (F, G, H) if f is not None)
        arrs, strides = [], []
        for d, f in enumerate((F, G, H)):
            if f is None:
                arrs.append(None)
                strides.append(None)
                continue
            want = [nv, n0, n1, n2]
            want[d + 1] += 1
            if list(f.shape) != want or f.dtype != first.dtype:
                raise ValueError(f"face fluxes of shape {tuple(f.shape)} / {f.dtype} for cells {tuple(shape)}")
            if f.stride(3) != 1:
                f = f.contiguous()
            arrs.append(f)
            strides.append((_lib.ctypes.c_longlong * 3)(*f.stride()[:3]))
        args = []
        for a, st in zip(arrs, strides):
            args += [_lib.ptr(a), st]
        return _lib.sfx(first), args, arrs

    def _euler_equation_into(self, F, G, H, out):
        """dudt into out from face fluxes that may be views into larger arrays (last axis contiguous)."""
        nv, n0, n1, n2 = out.shape
        active = sum(1 << d for d in range(3) if self.active[d])
        sfx, args, _keep = self._flux_args(F, G, H, out.shape)
        _lib.call("fvhoe_euler_equation_strided" + sfx, *args, _lib.ptr(out), nv, n0, n1, n2, active,
                  -(1 / self.h[0]), -(1 / self.h[1]), -(1 / self.h[2]), _lib.stream())

    def euler_equation(self, F, G, H):
        """dudt from face fluxes (fvhoe/solver.py:734-751)."""
        out = _lib.empty((self.nvars, *self.n))
        self._euler_equation_into(F, G, H, out)
        return out

    # ------------------------------------------------------------ bookkeeping
    def log_troubles(self, trouble=None, NAD_violation_magnitude=None, PAD_violation_magnitude=None,
                     reset: bool = False, gate=None):
        """Per-step means of trouble and violation magnitudes (fvhoe/solver.py:794-833);
        the counter starts at 1 (fvhoe/solver.py:400).  gate (device scalar, the step's dt): the
        device work does nothing when it is 0, a step that stays at its target time."""
        lg = self.log
        n = self._trouble_count.numel()

        def log(dst, src, mode, d):
            if gate is None:
                _lib.call("fvhoe_log_troubles", _lib.ptr(dst), _lib.ptr(src), n, mode, d, _lib.stream())
            else:
                _lib.call("fvhoe_log_troubles_gated", _lib.ptr(dst), _lib.ptr(src), n, mode, d, _lib.ptr(gate),
                          _lib.stream())

        gp = _lib.ptr(gate) if gate is not None else None
        if reset:
            if self.trouble_counter > 0:
                # the trouble flags were counted since the last reset (exact integer sums); the mean is
                # written from the counts here
                _lib.call("fvhoe_log_mean", _lib.ptr(dict.__getitem__(lg, "mean trouble")),
                          _lib.ptr(self._trouble_count), n, float(self.trouble_counter), gp, _lib.stream())
                self._count_pending = False
                for k in lg:
                    if k == "mean trouble" or not self._log_written.get(k, False):
                        continue  # never written since construction: dividing zeros changes nothing
                    log(lg[k], None, 2, float(self.trouble_counter))
                if self._mag is not None:  # the fused stage's magnitudes: the mean is formed when read
                    _lib.call("fvhoe_set_scalar_gated", _lib.ptr(self._mag["div"]), float(self.trouble_counter), gp,
                              _lib.stream())
                    self._mag_closed = True
                self.trouble_counter = 0
            return
        t8 = trouble if trouble.dtype == _lib.torch().int8 else trouble.to(_lib.torch().int8).contiguous()
        _lib.call("fvhoe_log_count", _lib.ptr(self._trouble_count), _lib.ptr(t8), n,
                  0 if self.trouble_counter == 0 else 1, gp, _lib.stream())
        self._count_pending = True
        for k, v in (("mean NAD mag", NAD_violation_magnitude), ("mean PAD mag", PAD_violation_magnitude)):
            if v is None:
                continue
            self._log_written[k] = True
            _lib.call("fvhoe_rk_combine", _lib.RK_AXPY, _lib.ptr(lg[k]),
                      _lib.ptr(self._zero_like_log if self.trouble_counter == 0 else lg[k]), _lib.ptr(v.contiguous()),
                      None, None, None, _lib.ptr(self._one), n, _lib.stream())
        self.trouble_counter += 1

    def _log_troubles_device(self, flags, gate=None):
        self.log_troubles(flags, None, None, gate=gate)

    def step_helper_function(self):
        if self.a_posteriori_slope_limiting:
            # gated by the step's dt: a step that stays at its target time leaves the log as it is
            self.log_troubles(reset=True, gate=self._scal[4:5] if self._fused_rk() else None)

    def snapshot(self):
        """fvhoe/solver.py:840-874."""
        w = self.interpolate_primitives_from_conservatives(self.u, self.p, fv_average=self.snapshots_as_fv_averages)
        rec = {"x": self.x, "y": self.y, "z": self.z, "t": self.t, "w": w.cpu().numpy()}
        if self.a_posteriori_slope_limiting:
            for key, name in (("mean trouble", "trouble"), ("mean NAD mag", "NAD violation magnitude"),
                              ("mean PAD mag", "PAD violation magnitude")):
                rec[name] = np.broadcast_to(self.log[key].cpu().numpy(), (5, *self.n)).copy()
        self.snapshots.append(rec)
        self.snapshot_times.append(self.t)
        if self.snapshot_helper_function is not None:
            self.snapshot_helper_function(self)

    def read_snapshots(self) -> bool:
        if os.path.exists(self.snapshot_dir):
            with open(os.path.join(self.snapshot_dir, "arrs.pkl"), "rb") as f:
                self.snapshots = pickle.load(f)
            return True
        return False

    def write_snapshots(self, overwrite: bool):
        """arrs.pkl, attrs.json, timing.txt (fvhoe/solver.py:889-953)."""
        d = self.snapshot_dir
        if os.path.exists(d) and not overwrite:
            raise FileExistsError(f"Snapshot directory {d} already exists.")
        if os.path.exists(d):
            shutil.rmtree(d)
        os.makedirs(d)
        with open(os.path.join(d, "arrs.pkl"), "wb") as f:
            pickle.dump(self.snapshots, f)
        with open(os.path.join(d, "attrs.json"), "w") as f:
            json.dump(self.snapshot_attrs(), f, indent=4)
        with open(os.path.join(d, "timing.txt"), "w") as f:
            f.write(timing_report(self.timer))

    def snapshot_attrs(self) -> dict:
        """The run's settings as written to attrs.json (fvhoe/solver.py:909-943): every attribute that JSON
        can hold (arrays, tensors, functions and the snapshots themselves left out), the initial condition
        by its function name, the boundary conditions, timer and device arrays as dictionaries, infinite PAD
        bounds as strings, keys sorted case-insensitively."""
        skip = ("progress_bar", "timestamps", "snapshots", "w0_cell_centers_cache")
        attrs = {}
        for k, v in self.__dict__.items():
            if k.startswith("_") or k in skip:
                continue
            if k == "w0":
                attrs[k] = getattr(v, "__name__", "__name__ not found")
            elif k == "bc":
                attrs[k] = v.to_dict()
            elif k == "timer":
                attrs[k] = {c: round(float(x), 2) for c, x in v.items()}
            elif k == "PAD":
                attrs[k] = {q: [{np.inf: "inf", -np.inf: "-inf"}.get(x, x) for x in b] for q, b in v.items()}
            elif _jsonable(v):
                attrs[k] = v
        # the device arrays the run holds (the reference's array manager lists its arrays by name)
        attrs["am"] = {"names": ["u", "mean trouble", "mean NAD mag", "mean PAD mag"]}
        attrs["commit_details"] = commit_details()
        return {key: attrs[key] for key in sorted(attrs, key=lambda x: (x.lower(), x))}

    def run(self, *args, **kwargs):
        """RK order matched to max(p), capped at RK4 (fvhoe/solver.py:955-974)."""
        q = min(max(self.p), 3)
        {0: self.euler, 1: self.ssprk2, 2: self.ssprk3, 3: self.rk4}[q](*args, **kwargs)
