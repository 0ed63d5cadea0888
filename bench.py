"""Benchmark: FVHOE per-RK-stage update on one MI355X.

Workload: periodic 2-D Kelvin-Helmholtz, 2048^2 cells, p=3, HLLC, SSP-RK3,
a-posteriori NAD/PAD/SED check with MUSCL fallback (gamma 1.4, CFL 0.8,
seeded IC).  A step is one SSP-RK3 step = 3 stages over the whole grid; the
state lives in HBM for the whole timed region and the host never synchronises
inside it.  Three forms are timed from the same initial state: the plain FP64
path (bit-identical to the reference), the FP32 bulk form (FP32 sweeps and
Riemann solves; FP64 pressure recovery, trouble tests and RK accumulation),
and the learned form (FP32 bulk plus the troubled-cell predictor, run once
per step, whose pre-flagged cells get MUSCL fluxes before the candidate; the
plain check still runs on every cell, the headline), also timed with the
network on a side stream beside each step's first stage ("learned-overlap":
no faster, since the stage's kernels already occupy every CU, and it slows the
candidate kernel it shares the chip with).

    python bench.py [--steps K] [--warmup W] [--n N] [--forms fp64,fp32,learned-overlap,learned] [--no-cpu]

Prints one JSON line.  `value` is Mcell-updates/s per RK stage (cells x
stages / s) of the last form listed.  `roofline` is for the kernel that takes
the most time, from HIP events around each launch on the stream it runs on;
`stage_roofline` uses the algorithmic bytes of a whole a-posteriori FP64
stage.  `cpu_baseline` times the NumPy oracle (one thread) on two stages of
the same workload.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BREAKDOWN_STEPS = 3  # eager steps timed kernel by kernel after the timed graph replays
# algorithmic bytes per cell for one launch of each fused-stage kernel (FP64, 5 variables)
KERNELS = ["prims", "flux_x", "flux_y", "prims_candidate", "detect", "fallback", "final", "predict"]
BYTES_PER_CELL = {
    "prims": 40 + 40,             # read u, write W
    "flux_x": 40 + 40,            # read W, write F
    "flux_y": 40 + 40,            # read W, write G
    "flux_xy": 40 + 80,           # 2-D: both face axes from one W box (read W, write F and G)
    # 2-D: the candidate tile also runs the troubled-cell test (NAD over W's 3x3, PAD) on its cells
    "prims_candidate": 120 + 40 + 40 + 1,  # read u, F, G, write WS; read W, write flags
    "detect": 0,                  # the SED pass over the cells whose NAD test flagged: scales with them only
    "fallback": 0,                # work and traffic scale with the flagged cells only
    "final": 80 + 40 + 80,        # read F, G, write dudt; RK combine reads u0 (+ k) and writes the next state
    "predict": 32 + 8 + 8 + 1,    # features: read W (rho, vx, vy, P), write 4 bf16; network: read them, write flags
}
# bytes of the arrays stored as the bulk type (W, F, G, WS) per cell for each kernel, FP64
BULK_BYTES = {"prims": 40, "flux_x": 80, "flux_y": 80, "flux_xy": 120, "prims_candidate": 160, "detect": 0,
              "fallback": 0,
              "final": 80, "predict": 32}
STAGE_BYTES_APOST_FP64 = 40 * (5 / 3 + 1) + 120  # 226.7 B per cell-stage
BF16_PEAK_TFLOPS = 2500.0
# useful multiply-adds of the predictor network per cell: 2 * 25 * (4*16 + 16*16 + 16*1)
PREDICT_FLOP_PER_CELL = 2 * 25 * (4 * 16 + 16 * 16 + 16)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "kh2048_hbm_traffic.json")


def measured_traffic(form, kernel):
    """HBM bytes per launch of `kernel` in `form` from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    recorded in profiles/ (null when that kernel was not measured)."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
        return rec["forms"][form][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--forms", default="fp64,fp32,learned-overlap,learned",
                    help="comma list of forms to time (fp64, fp32, learned, learned-overlap = the network on a "
                         "side stream beside each step's first stage); the last one is the headline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=2048)
    ap.add_argument("--cpu-stages", type=int, default=2)
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run 3 device copies of 167,772,160 bytes (scale for HBM byte counters)")
    ap.add_argument("--workload", default="kh", choices=["kh", "ensemble", "kh-accuracy", "predictor", "wave3d"],
                    help="kh: the headline Kelvin-Helmholtz stage; ensemble: 512 uneven 1-D Shu-Osher runs to t=1.8; "
                         "kh-accuracy: every form to --t-end against a 2x finer FP64 reference; "
                         "predictor: the inference kernel alone on random features; "
                         "wave3d: the 3-D density wave at --n3^3, p=7, LLF, no limiter")
    ap.add_argument("--n3", type=int, default=256)
    ap.add_argument("--cpu-n3", type=int, default=176, help="wave3d: grid of the one-stage CPU baseline sample")
    ap.add_argument("--wave-t", type=float, default=None,
                    help="wave3d: after the timed steps run each form on to this time and compare the final states")
    ap.add_argument("--t-end", type=float, default=0.8)
    ap.add_argument("--ensemble-t", type=float, default=1.8)
    ap.add_argument("--ensemble-count", type=int, default=512)
    return ap.parse_args()


def run_kh_accuracy(a):
    """L1 density error at t_end of each form at n^2 against plain FP64 at (2n)^2 started from the
    piecewise-constant refinement of the same initial averages and averaged back over 2x2 blocks."""
    import numpy as np
    import torch

    from fvhoe_amd import initial_conditions as ics
    from fvhoe_amd.boundary_conditions import BoundaryCondition
    from fvhoe_amd.predictor import DEFAULT_WEIGHTS, TroubledCellPredictor
    from fvhoe_amd.solver import EulerSolver

    n = a.n
    u0 = ics.kelvin_helmholtz(n, seed=0)

    def solver(u, m, form):
        class IC(ics.InitialCondition):
            def base_ic(self):
                return lambda x, y, z: u

        pred = TroubledCellPredictor.load(DEFAULT_WEIGHTS) if form.startswith("learned") else None
        s = EulerSolver(w0=IC(), nx=m, ny=m, px=3, py=3, gamma=1.4, CFL=0.8, conservative_ic=True, fv_ic=True,
                        bc=BoundaryCondition(), riemann_solver="hllc", a_posteriori_slope_limiting=True,
                        progress_bar=False, bulk_precision="fp64" if form == "fp64" else "fp32", predictor=pred,
                        predictor_overlap=form == "learned-overlap")
        s.integrator = "ssprk3"
        return s

    def run(s):
        t0 = time.perf_counter()
        steps = 0
        while s.t < a.t_end:
            s.take_step(a.t_end)
            steps += 1
            if steps % 500 == 0:
                print(f"#   step {steps} t={s.t:.4f}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        return steps, time.perf_counter() - t0

    fine_u0 = np.repeat(np.repeat(u0, 2, axis=1), 2, axis=2)
    print(f"# reference {2 * n}^2 fp64 to t={a.t_end}", file=sys.stderr, flush=True)
    ref = solver(fine_u0, 2 * n, "fp64")
    rsteps, rsec = run(ref)
    r = ref.u[0, :, :, 0]
    rho_ref = 0.25 * (r[0::2, 0::2] + r[1::2, 0::2] + r[0::2, 1::2] + r[1::2, 1::2])
    del ref
    out = {"workload": f"kelvin_helmholtz_{n}x{n}_p3_hllc_aposteriori_ssprk3_to_t{a.t_end}",
           "reference": {"cells": (2 * n) ** 2, "steps": rsteps, "seconds": round(rsec, 2)}, "forms": {}}
    dens = {}
    for form in a.forms.split(","):
        print(f"# {form} {n}^2 to t={a.t_end}", file=sys.stderr, flush=True)
        s = solver(u0, n, form)
        steps, sec = run(s)
        rho = s.u[0, :, :, 0]
        dens[form] = rho.clone()
        rec = {"steps": steps, "seconds": round(sec, 2), "t": s.t,
               "l1_density_error": float((rho - rho_ref).abs().mean().item())}
        if s.predictor_stats is not None:
            rec["predictor"] = dict(zip(("preflagged", "second_pass", "preflag_unflagged"),
                                        [int(v) for v in s.predictor_stats.tolist()]))
        out["forms"][form] = rec
        del s
    if "fp64" in out["forms"]:
        e64 = out["forms"]["fp64"]["l1_density_error"]
        for f, rec in out["forms"].items():
            rec["error_ratio_vs_fp64"] = round(rec["l1_density_error"] / e64, 4)
            rec["l1_density_vs_fp64"] = float((dens[f] - dens["fp64"]).abs().mean().item())
    print(json.dumps(out), flush=True)


def run_predictor(a):
    """The predictor's inference kernel alone: n^2 logits from uniform random planar features,
    `steps` timed launches after `warmup`."""
    import torch

    from fvhoe_amd import _lib
    from fvhoe_amd.predictor import DEFAULT_WEIGHTS, TroubledCellPredictor

    n = a.n
    m = TroubledCellPredictor.load(DEFAULT_WEIGHTS)
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    X = torch.rand((4, n + 12, n + 12), dtype=torch.float32, device="cuda", generator=g)
    flags = torch.empty((n, n), dtype=torch.int8, device="cuda")

    def launch():
        _lib.call("fvhoe_pred_infer", _lib.ptr(X), n, n, _lib.ptr(m.packed), float(m.logit_threshold),
                  _lib.ptr(flags), None, _lib.stream())

    for _ in range(a.warmup):
        launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.steps):
        launch()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    tf = PREDICT_FLOP_PER_CELL * n * n / (ms / 1e3) / 1e12
    print(json.dumps({"metric": "predictor inference", "value": round(tf, 1), "unit": "TFLOP/s", "ms": round(ms, 4),
                      "cells": n * n, "frac_bf16_peak": round(tf / BF16_PEAK_TFLOPS, 4)}), flush=True)


def run_wave3d(a):
    """The 3-D density wave rho = 1 + 0.5 sin 2 pi (x+y+z), v = (1,1,1), P = 1 on the periodic unit cube at
    n^3, p=7, LLF, SSP-RK3, no limiter (composed path), initial state built on the device, for each of the
    forms fp64 (bit-identical to the reference) and fp32 (float32 intermediates): `steps` timed
    host-synchronised steps after one warm-up step, and the L1 density error against the exact translated
    wave.  With --wave-t T each form then runs on to t = T (the last step clipped to land on T) and the
    L1 density errors at T against the exact wave and between the forms are reported."""
    import torch

    from fvhoe_amd import initial_conditions as ics
    from fvhoe_amd.solver import EulerSolver

    n = a.n3
    wave = ics.Sinus(dims="xyz", rho_min_max=(0.5, 1.5), P=1, vx=1, vy=1, vz=1)
    X, Y, Z = ics.device_mesh((n, n, n))
    h = (1.0 / n,) * 3

    def exact_rho(t):
        ic = wave.build_ic()
        return ics.fv_average(lambda x, y, z: ics.primitives_to_conservatives(ic(x - t, y - t, z - t), 1.4),
                              X, Y, Z, h, (7, 7, 7))[0]

    forms = [f for f in a.forms.split(",") if f in ("fp64", "fp32")] or ["fp64"]
    finals = {}
    for form in forms:
        t0 = time.perf_counter()
        s = EulerSolver(w0=wave, nx=n, ny=n, nz=n, px=7, py=7, pz=7, gamma=1.4, riemann_solver="llf",
                        a_posteriori_slope_limiting=False, progress_bar=False, ic_device=True, bulk_precision=form)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        s.integrator = "ssprk3"
        for _ in range(max(a.warmup, 1)):
            s.take_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            s.take_step()
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        ms = sec / a.steps * 1e3
        l1 = float((s.u[0] - exact_rho(s.t)).abs().mean().item())
        value = n ** 3 * 3 / (ms / 1e3) / 1e6
        # compulsory state traffic per cell-stage, SSP-RK3 minimum (5/3 state reads, 1 write): 106.7 B; the
        # state and the RK accumulation stay FP64 in both forms
        bytes_alg = 40 * (5 / 3 + 1)
        gbs = bytes_alg * value * 1e6 / 1e9
        line = {"metric": "Mcell-updates/s per RK stage", "value": round(value, 1), "unit": "Mcell-updates/s",
                "ms_per_step": round(ms, 2), "steps": a.steps, "dtype": "f64" if form == "fp64" else "f32",
                "config": {"workload": f"density_wave_{n}^3_p7_llf_ssprk3_nolimiter", "path": "composed",
                           "form": form},
                "setup_s": round(setup, 2), "t": s.t, "l1_density_vs_exact": l1,
                "stage_roofline": {"bytes_per_cell_stage": round(bytes_alg, 1), "achieved": round(gbs, 1),
                                   "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}}
        if form == forms[0] and not a.no_cpu:
            line["cpu_baseline"] = cpu_baseline_wave3d(a.cpu_n3)
        if a.wave_t:
            t0 = time.perf_counter()
            steps = 0
            while s.t < a.wave_t:
                s.take_step(target_time=a.wave_t)
                steps += 1
                if steps % 200 == 0:
                    print(f"# wave3d {form}: t = {s.t:.4f} after {steps} steps", file=sys.stderr, flush=True)
            torch.cuda.synchronize()
            line["run_to_t"] = {"t": s.t, "steps": steps, "seconds": round(time.perf_counter() - t0, 1),
                                "l1_density_vs_exact": float((s.u[0] - exact_rho(s.t)).abs().mean().item())}
            finals[form] = s.u[0].clone()
        print(json.dumps(line), flush=True)
        del s
        torch.cuda.empty_cache()
    if len(finals) == 2:
        d = float((finals["fp32"] - finals["fp64"]).abs().mean().item())
        print(json.dumps({"metric": "wave3d L1 density difference fp32 vs fp64", "value": d, "t": a.wave_t,
                          "n": n}), flush=True)


def run_ensemble(a):
    """512 Shu-Osher-type 1-D instances, N ~ U{256..8192}, p=5 HLLC a-posteriori, SSP-RK3 to t=1.8, "ic"
    boundaries, every step one captured graph over all instances; tile size and stage form (chain of
    launches or one persistent kernel per stage) from the autotuner.  Forms: fp64, fp32 and learned (the
    FP32 form with the troubled-cell predictor on the runs seen as planar 2-D states, the chain of
    launches); with learned, an untimed FP32 run counts the plain check's flags for the second-pass
    share, and the final densities of the two are compared."""
    import numpy as np
    import torch

    from fvhoe_amd.ensemble import shu_osher_ensemble_solver
    from fvhoe_amd.predictor import TroubledCellPredictor

    forms, finals = {}, {}
    names = []
    for form in a.forms.split(","):
        f = "learned" if form.startswith("learned") else form
        if f not in names:
            names.append(f)
    for form in names:
        if form == "learned":
            e = shu_osher_ensemble_solver(a.ensemble_count, seed=0, bulk_precision="fp32",
                                          predictor=TroubledCellPredictor.load())
        else:
            e = shu_osher_ensemble_solver(a.ensemble_count, seed=0, bulk_precision=form)
        tune = {f"{k[0]}/{k[1]}": round(v, 4) for k, v in e.autotune(a.ensemble_t, steps=10).items()}
        print(f"# ensemble {form}: {sum(e.n)} cells, autotune {tune} -> tile {e.tile}, "
              f"{'persistent' if e.persistent else 'chain'}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        steps = e.run(a.ensemble_t)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        cu = e.cell_updates()
        forms[form] = {"value": round(cu / wall / 1e6, 2), "seconds": round(wall, 3), "ensemble_steps": steps,
                       "cell_updates": cu, "tile": e.tile, "stage_form": "persistent" if e.persistent else "chain",
                       "autotune_ms_per_step": tune,
                       "t_min": float(e.times().min()), "t_max": float(e.times().max())}
        if form == "learned":
            st = e.predictor_stats.cpu().numpy()
            forms[form]["predictor"] = {"preflagged": int(st[0]), "second_pass": int(st[1]),
                                        "preflag_unflagged": int(st[2]),
                                        "logit_threshold": e.predictor.logit_threshold}
        if form in ("fp32", "learned"):
            finals[form] = [e.state(i)[0] for i in range(e.count)]
        print(f"# ensemble {form}: {wall:.2f}s, {steps} steps", file=sys.stderr, flush=True)
        del e
        torch.cuda.empty_cache()
    if "learned" in forms:
        # the plain FP32 check's flags over the same run (untimed), and the learned form's density
        # difference from the plain FP32 run at the end time
        c = shu_osher_ensemble_solver(a.ensemble_count, seed=0, bulk_precision="fp32", count_troubles=True)
        c.run(a.ensemble_t)
        plain_flags = int(c.predictor_stats[1].item())
        if "fp32" not in finals:
            finals["fp32"] = [c.state(i)[0] for i in range(c.count)]
        del c
        pr = forms["learned"]["predictor"]
        pr["check_flags_without_predictor"] = plain_flags
        pr["second_pass_share"] = round(pr["second_pass"] / max(plain_flags, 1), 4)
        pr["preflag_per_check_flag"] = round(pr["preflagged"] / max(plain_flags, 1), 4)
        d = [float(np.abs(x - y).mean()) for x, y in zip(finals["learned"], finals["fp32"])]
        forms["learned"]["l1_density_vs_fp32"] = {"mean": float(np.mean(d)), "max": float(np.max(d))}
    head = forms[list(forms)[-1]]
    out = {"metric": "Mcell-updates/s per RK stage (ensemble)", "value": head["value"], "unit": "Mcell-updates/s",
           "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if list(forms)[-1] == "fp64" else "f32", "data": "synthetic",
           "config": {"workload": f"shu_osher_ensemble_{a.ensemble_count}_p5_hllc_aposteriori_ssprk3_t{a.ensemble_t}",
                      "boundary": "ic", "form": list(forms)[-1], "parallelism": "single GPU"},
           "forms": forms}
    print(json.dumps(out), flush=True)


ONE_THREAD = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_child(kind: str, n: int, stages: int) -> None:
    """Runs in a child process started with the thread variables at 1 before NumPy loads: pins itself to
    one core, times the NumPy oracle and prints one JSON line."""
    cores = sorted(os.sched_getaffinity(0))
    core = cores[0]
    os.sched_setaffinity(0, {core})
    assert all(os.environ.get(k) == "1" for k in ONE_THREAD), "thread variables must be 1 before NumPy loads"
    import numpy as np

    import oracle as O

    if kind == "kh":
        from fvhoe_amd import initial_conditions as ics

        u = ics.kelvin_helmholtz(n, seed=0)
        s = O.Scheme(n=(n, n, 1), p=(3, 3, 0), gamma=1.4, a_posteriori=True)
        cells, what = n * n * stages, f"{stages} a-posteriori FP64 stages of Kelvin-Helmholtz {n}^2 p=3 HLLC"
    else:
        X, Y, Z = O.fv_uniform_meshgen((n, n, n))
        one = np.ones_like(X)
        u = O.compute_conservatives(np.stack([1 + 0.5 * np.sin(2 * np.pi * (X + Y + Z)), one, one, one, one]), 1.4)
        s = O.Scheme(n=(n, n, n), p=(7, 7, 7), gamma=1.4, riemann_solver="llf")
        stages = 1
        cells, what = n ** 3, (f"1 FP64 stage of the density wave at {n}^3 p=7 LLF (the cell-centre sample of the "
                               f"point-value IC)")
    t0 = time.perf_counter()
    for _ in range(stages):
        O.hydrodynamics(s, u)
    dt = time.perf_counter() - t0
    print(json.dumps({"value": round(cells / dt / 1e6, 4), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
                      "sample": f"{what} (NumPy oracle), {dt:.1f} s on core {core}; {len(cores)} cores in the "
                                f"process's affinity set, {os.cpu_count()} visible, 1 used",
                      "affinity": [core], "cpu_model": cpu_model(), "seconds": round(dt, 2),
                      "threads": {k: os.environ.get(k) for k in ONE_THREAD}, "numpy": np.__version__}), flush=True)


def _cpu_baseline_run(kind: str, n: int, stages: int):
    """The CPU baseline in a child process: one thread (OMP / OpenBLAS / MKL at 1 before NumPy loads),
    pinned to one core."""
    import subprocess

    env = dict(os.environ, **{k: "1" for k in ONE_THREAD})
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", kind, str(n), str(stages)],
                       env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(n: int, stages: int):
    """NumPy oracle, `stages` stages of KH n^2 on one pinned core."""
    return _cpu_baseline_run("kh", n, stages)


def cpu_baseline_wave3d(n: int):
    """NumPy oracle, one FP64 stage of the 3-D density wave at n^3 (p=7, LLF, no limiter) on one pinned core."""
    return _cpu_baseline_run("wave3d", n, 1)


def run_form(n, form, steps, warmup, u0):
    """Time `steps` SSP-RK3 steps of one form ("fp64", "fp32", "learned" = FP32 bulk with the
    troubled-cell predictor, "learned-overlap" = the same with the network beside each step's first
    stage); returns (result dict, final state)."""
    import torch

    from fvhoe_amd import _lib
    from fvhoe_amd import initial_conditions as ics
    from fvhoe_amd.boundary_conditions import BoundaryCondition
    from fvhoe_amd.predictor import DEFAULT_WEIGHTS, TroubledCellPredictor
    from fvhoe_amd.solver import EulerSolver

    class KH(ics.InitialCondition):
        def base_ic(self):
            return lambda x, y, z: u0

    pred = TroubledCellPredictor.load(DEFAULT_WEIGHTS) if form.startswith("learned") else None
    s = EulerSolver(w0=KH(), nx=n, ny=n, px=3, py=3, gamma=1.4, CFL=0.8, conservative_ic=True, fv_ic=True,
                    bc=BoundaryCondition(), riemann_solver="hllc", a_posteriori_slope_limiting=True,
                    progress_bar=False, bulk_precision="fp64" if form == "fp64" else "fp32", predictor=pred,
                    predictor_overlap=form == "learned-overlap")
    assert s.fused
    s.integrator = "ssprk3"
    print(f"# {form}: warmup {warmup} steps at {n}^2", file=sys.stderr, flush=True)
    s.advance(max(warmup, 2), graph=True)  # the first step runs eagerly, then one step is captured in a graph
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    s.advance(steps, graph=True)  # graph replays only
    stop.record()
    torch.cuda.synchronize()
    ms_total = start.elapsed_time(stop)
    # per-kernel breakdown: the same kernels launched eagerly, each bracketed by HIP events on its stream
    L = _lib.lib()
    _lib.check(L.fvhoe_stage_timing(1), "timing")
    s.advance(BREAKDOWN_STEPS)
    torch.cuda.synchronize()
    ms = (ctypes.c_double * 8)()
    cnt = (ctypes.c_longlong * 8)()
    _lib.check(L.fvhoe_stage_timing_read(ms, cnt), "timing read")
    _lib.check(L.fvhoe_stage_timing(0), "timing")
    cells = n * n
    stages = 3 * steps
    per_kernel = {}
    for i, k in enumerate(KERNELS):
        if cnt[i]:
            if k == "flux_x" and not cnt[KERNELS.index("flux_y")]:
                k = "flux_xy"  # the 2-D stage times its one flux kernel for both axes under the x slot
            avg = ms[i] / cnt[i]
            b = BYTES_PER_CELL[k] - (BULK_BYTES[k] / 2 if form != "fp64" else 0)
            per_kernel[k] = {"ms_avg": avg, "launches": cnt[i], "bytes_per_cell": b,
                             "algorithmic_GBs": b * cells / (avg / 1e3) / 1e9}
    print(f"# {form}: {ms_total / steps:.3f} ms/step", file=sys.stderr, flush=True)
    out = {"value": cells * stages / (ms_total / 1e3) / 1e6, "ms_per_step": ms_total / steps,
           "kernels": per_kernel, "t": s.t}
    if pred is not None:
        st = [int(v) for v in s.predictor_stats.tolist()]
        out["predictor"] = {"preflagged": st[0], "second_pass": st[1], "preflag_unflagged": st[2],
                            "logit_threshold": pred.logit_threshold}
    return out, s.u


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; sharding over several GPUs (x-slab domain "
                                   "decomposition with halo exchanges) is not built", "n_gpus": a.gpus}))
        sys.exit(2)
    if a.workload == "ensemble":
        run_ensemble(a)
        return
    if a.workload == "kh-accuracy":
        run_kh_accuracy(a)
        return
    if a.workload == "predictor":
        run_predictor(a)
        return
    if a.workload == "wave3d":
        run_wave3d(a)
        return
    import torch

    from fvhoe_amd import initial_conditions as ics

    n = a.n
    u0 = ics.kelvin_helmholtz(n, seed=0)
    if a.copy_probe:
        # a plain device copy of 167,772,160 bytes, three times: the scale for HBM byte counters
        src = torch.ones(5 * 2048 * 2048, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    names = a.forms.split(",")
    forms, states = {}, {}
    for form in names:
        forms[form], states[form] = run_form(n, form, a.steps, a.warmup, u0)
    head = forms[names[-1]]
    cells = n * n
    stages = 3 * a.steps
    pk = head["kernels"]
    # the dominant HBM-bound stage kernel: most time per step (average launch time x launches); the
    # network is MFMA-bound and gets its own roofline, timed where it runs alone on the stage's
    # stream (the learned form) when that form was run, since beside the stage it shares the chip
    main = {k: v for k, v in pk.items() if k != "predict"}
    dom = max(main, key=lambda k: main[k]["ms_avg"] * main[k]["launches"])
    achieved = pk[dom]["algorithmic_GBs"]
    traffic = measured_traffic(names[-1], dom)
    pred_roof = None
    pk_pred = forms["learned"]["kernels"] if "learned" in forms else pk
    if "predict" in pk_pred:
        tf = PREDICT_FLOP_PER_CELL * cells / (pk_pred["predict"]["ms_avg"] / 1e3) / 1e12
        pred_roof = {"bound": "mfma", "kernel": "predict", "achieved": round(tf, 1), "peak": BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(tf / BF16_PEAK_TFLOPS, 4),
                     "flop_per_cell": PREDICT_FLOP_PER_CELL}
    stage_gbs = STAGE_BYTES_APOST_FP64 * head["value"] * 1e6 / 1e9
    extra = {}
    for f in ("fp32", "learned", "learned-overlap"):
        if "fp64" in states and f in states:
            extra[f"l1_density_{f}_vs_fp64"] = (states[f][0] - states["fp64"][0]).abs().mean().item()
            extra[f"speedup_{f}_vs_fp64"] = round(forms[f]["value"] / forms["fp64"]["value"], 3)
    for f in forms.values():
        for v in f["kernels"].values():
            v["ms_avg"] = round(v["ms_avg"], 4)
            v["algorithmic_GBs"] = round(v["algorithmic_GBs"], 1)
        f["value"] = round(f["value"], 2)
        f["ms_per_step"] = round(f["ms_per_step"], 4)
    out = {
        "metric": "Mcell-updates/s per RK stage, learned+FP32 vs plain FP64, at matched L1 error",
        "value": head["value"],
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if names[-1] == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": f"kelvin_helmholtz_{n}x{n}_p3_hllc_aposteriori_ssprk3", "form": names[-1],
                   "cells": cells, "stages_timed": stages, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_cell": pk[dom]["bytes_per_cell"]},
        "predictor_roofline": pred_roof,
        "stage_roofline": {"bytes_per_cell_stage": round(STAGE_BYTES_APOST_FP64, 1),
                           "achieved": round(stage_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(stage_gbs / HBM_PEAK_GBS, 4)},
        "forms": forms,
        **extra,
    }
    if not a.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(a.cpu_n, a.cpu_stages)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    if len(sys.argv) == 5 and sys.argv[1] == "--cpu-baseline-child":
        cpu_baseline_child(sys.argv[2], int(sys.argv[3]), int(sys.argv[4]))
    else:
        main()
