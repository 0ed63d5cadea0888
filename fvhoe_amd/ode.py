"""Time integration with device-resident state (fvhoe/ode.py).

The steppers keep the reference's k-form arithmetic (fvhoe/ode.py:245-298) in
one HIP kernel per combination; dt never leaves the device inside a step.
The time loop reads the step's dt back once per step (the reference syncs
twice per stage), or not at all for a fixed step count.
"""

from __future__ import annotations

import time
from abc import ABC, abstractmethod
from typing import Any, Optional

import numpy as np

from fvhoe_amd import _lib


def _hit_target(t: float, dt: float, target: float = None) -> float:
    """fvhoe/ode.py:11-22."""
    if target is not None and target - t < dt:
        return target - t
    return dt


class ODE(ABC):
    """u' = f(t, u) with u a device tensor (fvhoe/ode.py:25-298)."""

    TIMES_CAP = 256  # most steps replayed between two host synchronisations in integrate

    def __init__(self, u0, progress_bar: bool = True, cupy: bool = False):
        # cupy: accepted for the reference's signature; the state is always a device tensor here
        self.t = 0.0
        self.timestamps = [0.0]
        self.step_count = 0
        self.u = _lib.as_device(u0)[0].clone()
        self.snapshots = []
        self.snapshot_times = []
        self.print_progress_bar = bool(progress_bar)
        self.timer = {"TOTAL": 0.0, "SNAPSHOTS": 0.0}
        n = self.u.numel()
        self._k = [_lib.empty(self.u.shape) for _ in range(4)]
        self._s = _lib.empty(self.u.shape)
        self._scal = _lib.zeros((8,))  # dt_stage x4, dt_step, t, target, spare
        self._n = n
        self._graphs = {}  # (integrator, t pointer, target) -> (captured step, RHS evaluations per step)
        # steps that moved since the last reset, and their end times (the first TIMES_CAP of them)
        self._steps_dev = _lib.zeros((1,), dtype=_lib.torch().int64)
        self._times_dev = _lib.zeros((self.TIMES_CAP,))

    def get_commit_details(self) -> dict:
        """Hash, author, date and branch of the package's git commit (fvhoe/ode.py:61-92)."""
        from fvhoe_amd.solver import commit_details

        return commit_details()

    def progress_bar_action(self, action: str, T: float = None):
        """Progress bar over simulated time: "setup", "update", "cleanup" (fvhoe/ode.py:201-216)."""
        if not self.print_progress_bar:
            return
        if action == "setup":
            from tqdm import tqdm

            self.progress_bar = tqdm(total=T, bar_format="{l_bar}{bar}| {n_fmt}/{total_fmt} [{elapsed}]")
        elif action == "update":
            self.progress_bar.n = self.t
            self.progress_bar.refresh()
        elif action == "cleanup":
            self.progress_bar.close()

    # ------------------------------------------------------------ to override
    @abstractmethod
    def stage(self, u, k_out, dt_out) -> None:
        """k_out = dudt(u); dt_out[0] = this stage's CFL dt (device)."""

    def f(self, t: float, u):
        """(dt, dudt) with dt a Python float (fvhoe/ode.py:234-243)."""
        k = _lib.empty(u.shape)
        dt = _lib.empty((1,))
        self.stage(u, k, dt)
        return float(dt.item()), k

    def snapshot(self):
        self.snapshots.append(None)
        self.snapshot_times.append(self.t)

    def step_helper_function(self):
        pass

    def _before_step(self) -> None:
        """Host work before a step that needs the host clock (time-dependent boundaries)."""

    def _needs_host_clock(self) -> bool:
        return False

    def read_snapshots(self) -> bool:
        return False

    def write_snapshots(self, overwrite: bool = False):
        pass

    # ---------------------------------------------------------------- steppers
    def _combine(self, kind, out, u, k0, k1=None, k2=None, k3=None, dt=None):
        _lib.call("fvhoe_rk_combine", kind, _lib.ptr(out), _lib.ptr(u), _lib.ptr(k0), _lib.ptr(k1), _lib.ptr(k2),
                  _lib.ptr(k3), _lib.ptr(dt), self._n, _lib.stream())

    def stage_rk(self, u, k_out, dt_out, kind, out, u0, k0=None, k1=None, k2=None, dt_step=None,
                 set_dt_step=False, t_dev=None, target=None, target_dev=None) -> bool:
        """Optional: a stage that also writes the RK combination (fused path).
        Returns False when not available."""
        return False

    def _device_step(self, target: Optional[float], t_dev, target_dev=None):
        """One step of self.integrator on self.u; returns the dt_step scalar view.  With target_dev
        (fused path) the target time is read from that device scalar when the step runs."""
        u, k, s = self.u, self._k, self._s
        sc = self._scal
        dts, dt_step = sc[0:1], sc[4:5]
        name = self.integrator
        first = dict(dt_step=dt_step, set_dt_step=True, t_dev=t_dev, target=target, target_dev=target_dev)
        if self._fused_rk():
            if name == "euler":
                self.stage_rk(u, k[0], dts, _lib.RK_AXPY, u, u, **first)
            elif name == "ssprk2":
                self.stage_rk(u, k[0], dts, _lib.RK_AXPY, s, u, **first)
                self.stage_rk(s, k[1], sc[1:2], _lib.RK_SSPRK2_FINAL, u, u, k0=s, dt_step=dt_step)
            elif name == "ssprk3":
                # k[1] receives k0 + k1, the sum the final combine starts with (same rounding)
                self.stage_rk(u, k[0], dts, _lib.RK_AXPY, s, u, **first)
                self.stage_rk(s, k[1], sc[1:2], _lib.RK_SSPRK3_STAGE3_SUM, s, u, k0=k[0], dt_step=dt_step)
                self.stage_rk(s, None, sc[2:3], _lib.RK_SSPRK3_FINAL_SUM, u, u, k0=k[1], dt_step=dt_step)
            elif name == "rk4":
                self.stage_rk(u, k[0], dts, _lib.RK_HALF_AXPY, s, u, **first)
                self.stage_rk(s, k[1], sc[1:2], _lib.RK_HALF_AXPY, s, u, dt_step=dt_step)
                self.stage_rk(s, k[2], sc[2:3], _lib.RK_AXPY, s, u, dt_step=dt_step)
                self.stage_rk(s, None, sc[3:4], _lib.RK_RK4_FINAL, u, u, k0=k[0], k1=k[1], k2=k[2],
                              dt_step=dt_step)
            else:
                raise ValueError(name)
            self._advance_clock(t_dev, dt_step)
            return dt_step
        self.stage(u, k[0], dts)
        _lib.call("fvhoe_hit_target", _lib.ptr(dt_step), _lib.ptr(dts), _lib.ptr(t_dev),
                  float(target) if target is not None else 0.0, int(target is not None), _lib.stream())
        if name == "euler":
            self._combine(_lib.RK_AXPY, u, u, k[0], dt=dt_step)
        elif name == "ssprk2":
            self._combine(_lib.RK_AXPY, s, u, k[0], dt=dt_step)
            self.stage(s, k[1], sc[1:2])
            self._combine(_lib.RK_SSPRK2_FINAL, u, u, s, k[1], dt=dt_step)
        elif name == "ssprk3":
            self._combine(_lib.RK_AXPY, s, u, k[0], dt=dt_step)
            self.stage(s, k[1], sc[1:2])
            self._combine(_lib.RK_SSPRK3_STAGE3, s, u, k[0], k[1], dt=dt_step)
            self.stage(s, k[2], sc[2:3])
            self._combine(_lib.RK_SSPRK3_FINAL, u, u, k[0], k[1], k[2], dt=dt_step)
        elif name == "rk4":
            self._combine(_lib.RK_HALF_AXPY, s, u, k[0], dt=dt_step)
            self.stage(s, k[1], sc[1:2])
            self._combine(_lib.RK_HALF_AXPY, s, u, k[1], dt=dt_step)
            self.stage(s, k[2], sc[2:3])
            self._combine(_lib.RK_AXPY, s, u, k[2], dt=dt_step)
            self.stage(s, k[3], sc[3:4])
            self._combine(_lib.RK_RK4_FINAL, u, u, k[0], k[1], k[2], k[3], dt=dt_step)
        else:
            raise ValueError(name)
        self._advance_clock(t_dev, dt_step)
        return dt_step

    def _advance_clock(self, t_dev, dt_step):
        _lib.call("fvhoe_advance_time_counted", _lib.ptr(t_dev), _lib.ptr(dt_step), _lib.ptr(self._steps_dev),
                  _lib.ptr(self._times_dev), self.TIMES_CAP, _lib.stream())

    def _fused_rk(self) -> bool:
        return False

    def take_step(self, target_time: float = None):
        """fvhoe/ode.py:183-193: one step, dt read back to advance the host clock."""
        t_dev = self._scal[5:6]
        _lib.call("fvhoe_set_scalar", _lib.ptr(t_dev), float(self.t), _lib.stream())
        self._before_step()
        dt = float(self._device_step(target_time, t_dev).item())
        self.t += dt
        self.timestamps.append(self.t)
        self.step_count += 1
        self.step_helper_function()

    def advance(self, n: int, graph: bool = False):
        """n steps with no host synchronisation (time kept on the device).  With graph=True the
        step after the first is captured once in a HIP graph (per integrator) and replayed, which
        removes the host launch work between kernels; the kernels and their arguments are the
        same as in the eager steps, so the results are identical."""
        if self._needs_host_clock():
            raise NotImplementedError("time-dependent boundary values need the host clock at every step: "
                                      "use take_step / integrate")
        t_dev = self._scal[5:6]
        _lib.call("fvhoe_set_scalar", _lib.ptr(t_dev), float(self.t), _lib.stream())
        # the capture needs the steady state of one completed step (workspace allocated, counters reset)
        have = (self.integrator, t_dev.data_ptr(), None) in self._graphs
        eager = n if not graph else (0 if have else min(n, 1))
        for _ in range(eager):
            self._device_step(None, t_dev)
            self.step_count += 1
            self.step_helper_function()
        if n > eager:
            g, per_step = self._step_graph(t_dev)
            for _ in range(n - eager):
                g.replay()
            self.step_count += n - eager
            if hasattr(self, "f_evaluation_count"):
                self.f_evaluation_count += per_step * (n - eager)
        self.t = float(t_dev.item())
        self.timestamps.append(self.t)

    def _step_graph(self, t_dev, target_dev=None):
        """One step of self.integrator (with its bookkeeping kernels) captured in a graph, toward the
        target time held in the device scalar target_dev if given.  The capture records the launches
        without running them; host-side counters touched while recording are restored, and they
        return to the same values after every step, so the captured launch arguments hold for every
        replay."""
        key = (self.integrator, t_dev.data_ptr(), None if target_dev is None else target_dev.data_ptr())
        if key not in self._graphs:
            torch = _lib.torch()
            fe = getattr(self, "f_evaluation_count", None)
            saved = {k: v for k, v in vars(self).items() if k in ("trouble_counter",)}
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            g = torch.cuda.CUDAGraph()
            with _lib.capture_guard(), torch.cuda.stream(s):
                with torch.cuda.graph(g, stream=s):
                    self._device_step(None if target_dev is None else 0.0, t_dev, target_dev)
                    self.step_helper_function()
            torch.cuda.current_stream().wait_stream(s)
            per_step = 0
            if fe is not None:
                per_step = self.f_evaluation_count - fe
                self.f_evaluation_count = fe
            for k, v in saved.items():
                setattr(self, k, v)
            self._graphs[key] = (_lib.GraphHolder(g), per_step)
        return self._graphs[key]

    def _chunkable(self) -> bool:
        """Steps can run back to back on the device: fused RK stages, no host clock needed."""
        return self._fused_rk() and not self._needs_host_clock()

    # steps between synchronisations as graph replays instead of eager launches: measured no faster
    # on MI355X (Sod 256: 186 vs 160 us/step, KH 256^2: 378-394 vs 360, KH 1024^2 equal within noise)
    CHUNK_GRAPHS = False

    def _run_steps(self, target: Optional[float], limit: int, graph: Optional[bool] = None) -> int:
        """Up to `limit` (<= TIMES_CAP) steps toward target, launched back to back (graph replays
        or eager launches) with one host synchronisation at the end; once a step lands on the target
        the remaining ones move nothing (dt = 0: state, time, trouble log and counters untouched).
        Updates t, timestamps, step_count and f_evaluation_count as take_step would have; returns
        the number of steps that moved."""
        limit = max(1, min(int(limit), self.TIMES_CAP))
        if graph is None:
            graph = self.CHUNK_GRAPHS
        t_dev = self._scal[5:6]
        _lib.call("fvhoe_set_scalar", _lib.ptr(t_dev), float(self.t), _lib.stream())
        target_dev = None
        if target is not None:  # one captured graph serves every target: it reads this scalar
            target_dev = self._scal[6:7]
            _lib.call("fvhoe_set_scalar", _lib.ptr(target_dev), float(target), _lib.stream())
        self._steps_dev.zero_()
        fe = getattr(self, "f_evaluation_count", None)
        eager = 0
        if self.step_count == 0:  # the capture needs one completed step (workspace allocated)
            self._device_step(target, t_dev)
            self.step_helper_function()
            eager = 1
        replays = limit - eager
        per_step = 0
        if replays > 0 and graph:
            g, per_step = self._step_graph(t_dev, target_dev)
            for _ in range(replays):
                g.replay()
        elif replays > 0:
            fe1 = getattr(self, "f_evaluation_count", None)
            for _ in range(replays):
                self._device_step(0.0 if target_dev is not None else None, t_dev, target_dev)
                self.step_helper_function()
            if fe1 is not None:  # counted per launched step on the host; keep the moving ones only
                per_step = (self.f_evaluation_count - fe1) // replays
                self.f_evaluation_count = fe1
        moved = int(self._steps_dev.item())
        times = self._times_dev[:moved].cpu().numpy()
        scal = self._scal.cpu().numpy()
        self.t = float(scal[5])
        self._dt_cfl = float(scal[0])  # the last step's CFL dt before any clipping to a target
        self.timestamps.extend(float(x) for x in times)
        self.step_count += moved
        if fe is not None:
            self.f_evaluation_count = self.f_evaluation_count + per_step * (moved - eager)
        return moved

    def _chunk_length(self, target: Optional[float]) -> int:
        """Steps to launch before the next synchronisation: enough to reach target at the last
        step's CFL dt (so few launches are wasted past it), at most TIMES_CAP."""
        if target is None:
            return self.TIMES_CAP
        dt = getattr(self, "_dt_cfl", 0.0)
        if not dt > 0.0:
            return 1
        return int(min(self.TIMES_CAP, max(1, np.ceil((target - self.t) / dt))))

    def integrate(self, T: float = None, n: int = None, exact: bool = True, downbeats: Any = None,
                  log_every_step: bool = False, snapshot_dir: str = None, overwrite: bool = False) -> None:
        """fvhoe/ode.py:94-174.  On the fused path the steps between snapshots run as graph replays in
        chunks, with the host clock read once per chunk instead of once per step."""
        if T is None:
            self.snapshot()
            t0 = time.time()
            if self._chunkable():
                left = n
                while left > 0:
                    left -= self._run_steps(None, min(left, self.TIMES_CAP))
            else:
                for _ in range(n):
                    self.take_step()
            _lib.torch().cuda.synchronize()
            self.timer["TOTAL"] += time.time() - t0
            self.snapshot()
            return
        save = snapshot_dir is not None
        if save:
            import os

            self.snapshot_dir = os.path.normpath(snapshot_dir)
            if not overwrite and self.read_snapshots():
                return
        if downbeats is None:
            downbeats = []
        elif isinstance(downbeats, np.ndarray):
            downbeats = downbeats.tolist()
        downbeats = sorted(set([0] + list(downbeats) + [T]))[1:]
        target = downbeats.pop(0) if exact else None
        chunked = target is not None and not log_every_step and self._chunkable()
        self.progress_bar_action(action="setup", T=T)
        self.snapshot()
        t0 = time.time()
        while self.t < T:
            if chunked:
                self._run_steps(target, self._chunk_length(target))
            else:
                self.take_step(target_time=target)
            self.progress_bar_action(action="update")
            if self.t == target or self.t >= T or log_every_step:
                ts = time.time()
                self.snapshot()
                self.timer["SNAPSHOTS"] += time.time() - ts
                if downbeats and self.t == target:
                    target = downbeats.pop(0)
        _lib.torch().cuda.synchronize()
        self.timer["TOTAL"] += time.time() - t0
        self.progress_bar_action(action="cleanup")
        if save:
            self.write_snapshots(overwrite)

    def euler(self, *args, **kwargs):
        self.integrator = "euler"
        self.integrate(*args, **kwargs)

    def ssprk2(self, *args, **kwargs):
        self.integrator = "ssprk2"
        self.integrate(*args, **kwargs)

    def ssprk3(self, *args, **kwargs):
        self.integrator = "ssprk3"
        self.integrate(*args, **kwargs)

    def rk4(self, *args, **kwargs):
        self.integrator = "rk4"
        self.integrate(*args, **kwargs)
