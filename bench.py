"""Benchmark: FVHOE per-RK-stage update on one MI355X.

Workload: periodic 2-D Kelvin-Helmholtz, 2048^2 cells, p=3, HLLC, SSP-RK3,
a-posteriori NAD/PAD/SED check with MUSCL fallback, FP64 (gamma 1.4, CFL 0.8,
seeded IC).  A step is one SSP-RK3 step = 3 stages over the whole grid; the
state lives in HBM for the whole timed region and the host never synchronises
inside it.

    python bench.py [--steps K] [--warmup W] [--n N] [--no-cpu]

Prints one JSON line.  `value` is Mcell-updates/s per RK stage (cells x
stages / s).  `roofline` is for the kernel that takes the most time,
from HIP events around each launch on the stream it runs on; `stage_roofline`
uses the algorithmic bytes of a whole a-posteriori FP64 stage.
`cpu_baseline` times the NumPy oracle (one thread) on one stage of the same
workload at 1024^2.
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
# algorithmic bytes per cell for one launch of each fused-stage kernel (FP64, 5 variables)
KERNELS = ["prims", "flux_x", "flux_y", "prims_candidate", "detect", "fallback", "final"]
BYTES_PER_CELL = {
    "prims": 40 + 40,             # read u, write W
    "flux_x": 40 + 40,            # read W, write F
    "flux_y": 40 + 40,            # read W, write G
    "prims_candidate": 120 + 40,  # read u, F, G, write WS
    "detect": 80 + 1,             # read W, WS, write flags
    "fallback": 0,                # work and traffic scale with the flagged cells only
    "final": 80 + 40,             # read F, G, write dudt
}
STAGE_BYTES_APOST_FP64 = 40 * (5 / 3 + 1) + 120  # 226.7 B per cell-stage


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=2048)
    ap.add_argument("--cpu-stages", type=int, default=2)
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run 3 device copies of 167,772,160 bytes (scale for HBM byte counters)")
    return ap.parse_args()


def cpu_baseline(n: int, stages: int):
    """NumPy oracle, `stages` stages of KH n^2 on one core."""
    import numpy as np

    import oracle as O
    from fvhoe_amd import initial_conditions as ics

    u = ics.kelvin_helmholtz(n, seed=0)
    s = O.Scheme(n=(n, n, 1), p=(3, 3, 0), gamma=1.4, a_posteriori=True)
    t0 = time.perf_counter()
    for _ in range(stages):
        O.hydrodynamics(s, u)
    dt = time.perf_counter() - t0
    return {"value": round(n * n * stages / dt / 1e6, 4), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{stages} a-posteriori FP64 stages of Kelvin-Helmholtz {n}^2 p=3 HLLC (NumPy oracle), "
                      f"{dt:.1f} s; {os.cpu_count()} logical cores visible, 1 used",
            "numpy": np.__version__}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md "
                                   "and not built", "n_gpus": a.gpus}))
        sys.exit(2)
    import torch

    from fvhoe_amd import _lib
    from fvhoe_amd import initial_conditions as ics
    from fvhoe_amd.boundary_conditions import BoundaryCondition
    from fvhoe_amd.solver import EulerSolver

    n = a.n
    u0 = ics.kelvin_helmholtz(n, seed=0)

    class KH(ics.InitialCondition):
        def base_ic(self):
            return lambda x, y, z: u0

    s = EulerSolver(w0=KH(), nx=n, ny=n, px=3, py=3, gamma=1.4, CFL=0.8, conservative_ic=True, fv_ic=True,
                    bc=BoundaryCondition(), riemann_solver="hllc", a_posteriori_slope_limiting=True,
                    progress_bar=False)
    assert s.fused
    s.integrator = "ssprk3"
    if a.copy_probe:
        # a plain device copy of 167,772,160 bytes, three times: the scale for HBM byte counters
        src = torch.ones(5 * 2048 * 2048, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    print(f"# warmup {a.warmup} steps at {n}^2", file=sys.stderr, flush=True)
    s.advance(a.warmup)
    torch.cuda.synchronize()
    L = _lib.lib()
    _lib.check(L.fvhoe_stage_timing(1), "timing")
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    s.advance(a.steps)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = start.elapsed_time(stop)
    ms = (ctypes.c_double * 7)()
    cnt = (ctypes.c_longlong * 7)()
    _lib.check(L.fvhoe_stage_timing_read(ms, cnt), "timing read")
    _lib.check(L.fvhoe_stage_timing(0), "timing")
    stages = 3 * a.steps
    cells = n * n
    value = cells * stages / (ms_total / 1e3) / 1e6
    per_kernel = {k: {"ms_avg": ms[i] / cnt[i], "launches": cnt[i]} for i, k in enumerate(KERNELS) if cnt[i]}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["ms_avg"])
    dom_bytes = BYTES_PER_CELL[dom] * cells
    achieved = dom_bytes / (per_kernel[dom]["ms_avg"] / 1e3) / 1e9
    for k, v in per_kernel.items():
        v["algorithmic_GBs"] = round(BYTES_PER_CELL[k] * cells / (v["ms_avg"] / 1e3) / 1e9, 1)
        v["ms_avg"] = round(v["ms_avg"], 4)
    stage_gbs = STAGE_BYTES_APOST_FP64 * cells * stages / (ms_total / 1e3) / 1e9
    out = {
        "metric": "Mcell-updates/s per RK stage, learned+FP32 vs plain FP64, at matched L1 error",
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_total / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"kelvin_helmholtz_{n}x{n}_p3_hllc_aposteriori_ssprk3", "form": "plain FP64",
                   "cells": cells, "stages_timed": stages, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_cell": BYTES_PER_CELL[dom]},
        "stage_roofline": {"bytes_per_cell_stage": round(STAGE_BYTES_APOST_FP64, 1),
                           "achieved": round(stage_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(stage_gbs / HBM_PEAK_GBS, 4)},
        "kernels": per_kernel,
        "wall_s": round(wall, 3),
    }
    if not a.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(a.cpu_n, a.cpu_stages)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
