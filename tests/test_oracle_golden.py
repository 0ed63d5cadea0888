"""The CPU oracle against vectors the reference itself wrote (tests/golden/make_golden.py).

FP64 results are compared bit for bit: the oracle follows the reference's
operation order, so any dropped term, swapped operand or reordered sum shows.
"""

import ast

import numpy as np
import pytest

import oracle as O
from cases import STAGE_CASES, build_stage_cases
from helpers import bits_equal, group, load, scheme_kwargs

build_stage_cases()
STAGES = load("stages.npz")
DET = load("detection.npz")
TRAJ = load("trajectories.npz")


@pytest.mark.parametrize("case", STAGE_CASES, ids=[c[0] for c in STAGE_CASES])
def test_stage_bitwise(case):
    name, n, p, bc, warm, seed, u0, kw = case
    g = group(STAGES, name)
    s = O.Scheme(**scheme_kwargs(n, p, bc, kw))
    log = O.StageLog()
    dt, dudt = O.hydrodynamics(s, g["u"], log)
    assert dt == float(g["dt"])
    assert bits_equal(dudt, g["dudt"]), np.abs(dudt - g["dudt"]).max()
    if "trouble" in g:
        assert np.array_equal(np.asarray(log.trouble).astype(np.int8), g["trouble"])
        assert bits_equal(log.NAD_mag, g["nad_mag"])
        assert bits_equal(log.PAD_mag, g["pad_mag"])


DET_KEYS = sorted({k.split("/")[0] for k in DET if k.startswith("d")}, key=lambda s: int(s[1:]))


@pytest.mark.parametrize("key", DET_KEYS)
def test_detect_troubled_cells(key):
    g = group(DET, key)
    meta = ast.literal_eval(str(g["meta"]).replace("inf", "1e999"))
    inp = group(DET, "in_" + meta["dims"])
    t, nm, pm = O.detect_troubled_cells(inp["w"], inp["ws"], meta["dims"], meta["eps"], meta["mode"],
                                        meta["range_type"], meta["NAD_vars"], meta["PAD"], meta["SED"], 1e-10)
    assert np.array_equal(np.asarray(t).astype(np.int8), g["trouble"])
    assert bits_equal(nm, g["nad_mag"])
    assert bits_equal(pm, g["pad_mag"])


@pytest.mark.parametrize("dims", ["x", "xy", "xz", "xyz"])
def test_smooth_extrema_detector(dims):
    g = group(DET, f"sed_{dims}")
    if len(dims) == 1:
        a = O.compute_1d_smooth_extrema_detector(g["u"], dims)
    elif len(dims) == 2:
        a = O.compute_2d_smooth_extrema_detector(g["u"], dims)
    else:
        a = O.compute_3d_smooth_extrema_detector(g["u"])
    assert bits_equal(a, g["alpha"])


@pytest.mark.parametrize("key", ["convex_x_0", "convex_x_1", "convex_xy_0", "convex_xy_1"])
def test_convex_broadcast(key):
    g = group(DET, key)
    dims = key.split("_")[1]
    per = bool(int(key.split("_")[2]))
    tx, ty, tz = O.broadcast_to_troubled_interfaces(g["trouble"], dims, True, (per, per, per))
    assert bits_equal(tx, g["tx"]) and bits_equal(ty, g["ty"]) and bits_equal(tz, g["tz"])


def _traj_scheme(name):
    P, Ob = ("periodic", "periodic"), ("outflow", "outflow")
    if name == "sod":
        return O.Scheme(n=(256, 1, 1), p=(3, 0, 0), bc=(Ob, P, P), gamma=1.4, a_posteriori=True)
    if name == "wave3d":
        return O.Scheme(n=(16, 16, 16), p=(7, 7, 7), gamma=1.4, riemann_solver="llf")
    if name == "quad64":
        return O.Scheme(n=(64, 64, 1), p=(3, 3, 0), bc=(Ob, Ob, P), gamma=1.4, a_posteriori=True)
    if name == "shuosher":
        return O.Scheme(n=(512, 1, 1), p=(5, 0, 0), bc=(Ob, P, P), domain=((-5, 5), (0, 1), (0, 1)),
                        gamma=1.4, a_posteriori=True)
    return O.Scheme(n=(64, 1, 1), p=(2, 0, 0), bc=(Ob, P, P), gamma=1.4, a_posteriori=True)


def test_sod_full_run():
    g = group(TRAJ, "sod")
    s = _traj_scheme("sod")
    t, u, steps = O.integrate(s, g["u0"], T=0.245)
    assert t == float(g["t"]) and steps == len(g["timestamps"]) - 1
    assert bits_equal(u, g["u_final"])
    assert bits_equal(O.primitive_snapshot(s, u), g["w_final"])


@pytest.mark.parametrize("name,nsteps,integ", [("wave3d", 2, "ssprk3"), ("quad64", 3, "ssprk3"),
                                               ("shuosher", 6, "ssprk3"), ("euler", 4, "euler"),
                                               ("ssprk2", 4, "ssprk2"), ("rk4", 4, "rk4")])
def test_trajectories(name, nsteps, integ):
    g = group(TRAJ, name)
    s = _traj_scheme(name)
    logs = [] if name == "quad64" else None
    t, u, _ = O.integrate(s, g["u0"], n=nsteps, integrator=integ, logs=logs)
    assert np.array_equal(np.cumsum(np.diff(g["timestamps"])) * 0 + g["timestamps"][1:], g["timestamps"][1:])
    assert t == float(g["timestamps"][-1])
    assert bits_equal(u, g["u_final"])
    if logs is not None:
        bits = np.stack([np.packbits(np.asarray(lg.trouble).astype(bool).ravel()) for lg in logs])
        assert np.array_equal(bits, g["trouble_bits"])
