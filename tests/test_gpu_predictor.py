"""Predictor kernels against the NumPy oracle (oracle/predictor_oracle.py, itself
pinned by finite differences): BF16-MFMA inference, FP32 training forward,
backward and Adam, the feature extraction of the fused stage, and what the
predicted stage does with its flags."""

import numpy as np
import pytest

from fvhoe_amd import _lib
from fvhoe_amd import initial_conditions as ics
from fvhoe_amd.predictor import Adam, TroubledCellPredictor, train
from gpu_helpers import assert_bits_equal, solver_from
from oracle import fvhoe_oracle as O
from oracle import predictor_oracle as PO

pytestmark = pytest.mark.gpu

P, Ob, R = ("periodic", "periodic"), ("outflow", "outflow"), ("reflective", "reflective")


def _theta(seed):
    th = PO.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    # non-zero biases so every term of the network is exercised
    for name, (o, n) in {"b1": (1600, 16), "b2": (1600 + 16 + 6400, 16), "b3": (PO.n_params() - 1, 1)}.items():
        th[o:o + n] = rng.uniform(-0.2, 0.2, n).astype(np.float32)
    return th


# 512 x 512 is 585 output tiles of 16 x 28: more tiles than the persistent grid's 512 blocks, so blocks
# loop over several tiles
@pytest.mark.parametrize("n0,n1", [(16, 16), (37, 23), (64, 48), (512, 512)])
def test_inference_matches_oracle(n0, n1):
    rng = np.random.default_rng(n0 * 100 + n1)
    th = _theta(n0)
    X = rng.random((4, n0 + 12, n1 + 12)).astype(np.float32)
    m = TroubledCellPredictor(th, logit_threshold=0.1)
    z = m.logits(X, n0, n1).cpu().numpy().astype(np.float64)
    zo, _ = PO.forward(th, X[None].astype(np.float64), rounding=True, act_bf16=True)
    zo = zo[0]
    err = np.abs(z - zo)
    tol = 2e-3 * np.abs(zo).max() + 1e-4
    print(f"[infer {n0}x{n1}] max|z| {np.abs(zo).max():.3f} max err {err.max():.2e} (tol {tol:.2e})", flush=True)
    assert err.max() <= tol
    # without the activation rounding the difference is the BF16 storage of activations only
    zf, _ = PO.forward(th, X[None].astype(np.float64), rounding=True, act_bf16=False)
    assert np.abs(z - zf[0]).max() <= 0.05 * np.abs(zf).max() + 1e-3
    f = m.flags(X, n0, n1).cpu().numpy()
    clear = np.abs(zo - 0.1) > tol
    assert np.array_equal(f[clear], (zo[clear] > 0.1).astype(np.int8))


def test_training_forward_and_backward_match_oracle():
    rng = np.random.default_rng(5)
    th = _theta(3)
    X = rng.standard_normal((2, 4, 22, 19)).astype(np.float32)
    y = (rng.random((2, 10, 7)) < 0.3).astype(np.int8)
    m = TroubledCellPredictor(th)
    z, cache = m.forward_train(X)
    loss, dz = m.loss_and_dlogits(z, y, 3.0)
    g = m.backward(dz, cache).cpu().numpy().astype(np.float64)
    zo, co = PO.forward(th, X.astype(np.float64), rounding=True)
    z = z.cpu().numpy().astype(np.float64)
    assert np.abs(z - zo).max() <= 1e-5 * (1 + np.abs(zo).max())
    lo = PO.loss(zo, y.astype(np.float64), 3.0)
    assert abs(float(loss.item()) - lo) <= 1e-5 * (1 + abs(lo))
    go = PO.backward(zo, co, y.astype(np.float64), 3.0)
    scale = np.abs(go).max()
    print(f"[grad] max|g| {scale:.3e} max err {np.abs(g - go).max():.2e}", flush=True)
    assert np.abs(g - go).max() <= 1e-4 * scale
    # per block of parameters as well (the small last layer would hide in the global scale)
    for a, b in ((0, 1600), (1600, 1616), (1616, 8016), (8016, 8032), (8032, 8432), (8432, 8433)):
        assert np.abs(g[a:b] - go[a:b]).max() <= 1e-4 * np.abs(go[a:b]).max() + 1e-9


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(2)
    th = rng.standard_normal(PO.n_params()).astype(np.float32)
    m = TroubledCellPredictor(th)
    grad = rng.standard_normal(PO.n_params()).astype(np.float32)
    opt = Adam(m, lr=0.01, eps=1e-8)
    opt.step(_lib.as_device(grad, _lib.torch().float32)[0])
    out = m.theta.cpu().numpy()
    # first step: bias-corrected moments are g and g^2, so the update is lr * g / (|g| + eps)
    ref = th - 0.01 * grad / (np.abs(grad) + 1e-8)
    assert np.abs(out - ref).max() <= 1e-6


def _quad_scheme(n, bc):
    return O.Scheme(n=(n, n, 1), p=(3, 3, 0), gamma=1.4, a_posteriori=True, bc=(bc, bc, P))


@pytest.mark.parametrize("bc", [Ob, P, R], ids=["outflow", "periodic", "reflective"])
@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_features_match_oracle(bc, form):
    n = 32
    u0 = ics.four_quadrants(n, seed=7)
    s = solver_from(u0, (n, n, 1), (3, 3, 0), (bc, bc, P), riemann_solver="hllc", a_posteriori_slope_limiting=True,
                    bulk_precision=form)
    k, dt = _lib.empty(s.u.shape), _lib.empty((1,))
    s.stage(s.u, k, dt)
    X = s.predictor_features().cpu().numpy()
    Xo = PO.features(_quad_scheme(n, bc), u0)
    if form == "fp64":
        assert np.array_equal(X.view(np.int32), Xo.view(np.int32))
    else:
        assert np.abs(X - Xo).max() <= 1e-5


def _stage_outputs(s):
    k, dt = _lib.empty(s.u.shape), _lib.empty((1,))
    s.stage(s.u, k, dt)
    return k.cpu().numpy(), dt.cpu().numpy(), s._trouble.cpu().numpy()


def test_predicted_stage_never_flagging_equals_plain():
    n = 48
    u0 = ics.four_quadrants(n, seed=3)
    kw = dict(riemann_solver="hllc", a_posteriori_slope_limiting=True)
    plain = solver_from(u0, (n, n, 1), (3, 3, 0), (Ob, Ob, P), **kw)
    never = TroubledCellPredictor(_theta(1), logit_threshold=1e30)
    pred = solver_from(u0, (n, n, 1), (3, 3, 0), (Ob, Ob, P), predictor=never, **kw)
    k0, dt0, t0 = _stage_outputs(plain)
    k1, dt1, t1 = _stage_outputs(pred)
    assert_bits_equal(k1, k0, "dudt")
    assert_bits_equal(dt1, dt0, "dt")
    assert np.array_equal(t1, t0)
    st = pred.predictor_stats.cpu().numpy()
    assert st[0] == 0 and st[1] == int(t0.sum()) and st[2] == 0


def test_predicted_stage_always_flagging_equals_forced_trouble():
    n = 48
    u0 = ics.four_quadrants(n, seed=4)
    kw = dict(riemann_solver="hllc", a_posteriori_slope_limiting=True)
    forced = solver_from(u0, (n, n, 1), (3, 3, 0), (Ob, Ob, P), force_trouble=True, **kw)
    always = TroubledCellPredictor(_theta(2), logit_threshold=-1e30)
    pred = solver_from(u0, (n, n, 1), (3, 3, 0), (Ob, Ob, P), predictor=always, **kw)
    k0, dt0, _ = _stage_outputs(forced)
    k1, dt1, t1 = _stage_outputs(pred)
    assert_bits_equal(k1, k0, "dudt")
    assert_bits_equal(dt1, dt0, "dt")
    assert t1.all()
    st = pred.predictor_stats.cpu().numpy()
    assert st[0] == n * n and st[1] == 0


def test_predicted_stage_flags_match_oracle_of_the_preblended_candidate():
    """With an arbitrary pre-flag map the stage is: MUSCL fluxes at the faces of
    pre-flagged cells, the check on that candidate, MUSCL at the faces of its
    flags too.  Rebuilt with oracle pieces and compared bit for bit."""
    n = 40
    u0 = ics.four_quadrants(n, seed=9)
    th = _theta(4)
    m = TroubledCellPredictor(th, logit_threshold=0.0)
    s = solver_from(u0, (n, n, 1), (3, 3, 0), (Ob, Ob, P), riemann_solver="hllc", a_posteriori_slope_limiting=True,
                    predictor=m)
    k, dt, trouble = _stage_outputs(s)
    # the pre-flags the stage used: the network on the oracle's features
    sch = _quad_scheme(n, Ob)
    X = PO.features(sch, u0)
    pre = m.flags(X, n, n).cpu().numpy().astype(bool)[:, :, None]
    assert 0 < pre.sum() < n * n
    dto, (F, G, H) = O.hydrofluxes(sch, u0, sch.p)
    _, (Fl, Gl, _) = O.hydrofluxes(sch, u0, (1, 1, 0), slope_limiter=sch.slope_limiter)
    tx, ty, _ = O.broadcast_to_troubled_interfaces(pre, sch.dims, False, (False, False, True))
    F1 = (1 - tx) * F + tx * Fl
    G1 = (1 - ty) * G + ty * Gl
    ustar = u0 + dto * O.euler_equation(sch, F1, G1, None)
    w = O.interpolate_primitives_from_conservatives(sch, u0, sch.p, (3, 3, 0))
    ws = O.interpolate_primitives_from_conservatives(sch, ustar, sch.p, (3, 3, 0))
    chk, _, _ = O.detect_troubled_cells(w[:5], ws[:5], sch.dims, sch.NAD, sch.NAD_mode, sch.NAD_range, sch.NAD_vars,
                                        sch.PAD, sch.SED, sch.SED_tolerance)
    allt = chk.astype(bool) | pre
    tx, ty, _ = O.broadcast_to_troubled_interfaces(allt, sch.dims, False, (False, False, True))
    F2 = (1 - tx) * F + tx * Fl
    G2 = (1 - ty) * G + ty * Gl
    dudt = O.euler_equation(sch, F2, G2, None)
    assert np.array_equal(trouble.astype(bool), allt)
    assert_bits_equal(dt[0], dto, "dt")
    assert_bits_equal(k, dudt, "dudt")
    st = s.predictor_stats.cpu().numpy()
    assert st[0] == pre.sum()
    assert st[1] == (chk.astype(bool) & ~pre).sum()
    assert st[2] == (pre & ~chk.astype(bool)).sum()


def _call_predicted(s, u, infer, overlap):
    k, dt = _lib.empty(s.u.shape), _lib.empty((1,))
    pa = s._predict_args(infer, overlap=s._overlap if overlap else None)
    _lib.call("fvhoe_stage_predicted", _lib.ctypes.byref(s._cfg), _lib.ptr(u), _lib.ptr(k), _lib.ptr(dt),
              _lib.ptr(s._trouble), s._workspace(), None, _lib.ctypes.byref(pa), _lib.stream())
    _lib.torch().cuda.synchronize()
    return k.cpu().numpy(), dt.cpu().numpy(), s._trouble.cpu().numpy()


@pytest.mark.parametrize("form", ["fp64", "fp32"])
@pytest.mark.parametrize("n", [48, 512])
def test_overlapped_inference(form, n):
    """A call that runs the network on the side stream is the plain stage bit for
    bit; the next call applies the same flags as the non-overlapped sequence (at 512^2 the side
    stream's 256 blocks each walk several of the 585 tiles)."""
    u0 = ics.four_quadrants(n, seed=5)
    kw = dict(riemann_solver="hllc", a_posteriori_slope_limiting=True, bulk_precision=form)
    m = TroubledCellPredictor(_theta(6), logit_threshold=0.0)
    plain = solver_from(u0, (n, n, 1), (3, 3, 0), (Ob, Ob, P), **kw)
    ov = solver_from(u0, (n, n, 1), (3, 3, 0), (Ob, Ob, P), predictor=m, predictor_overlap=True, **kw)
    seq = solver_from(u0, (n, n, 1), (3, 3, 0), (Ob, Ob, P), predictor=m, **kw)
    assert ov._overlap is not None
    k0, dt0, t0 = _stage_outputs(plain)
    k1, dt1, t1 = _call_predicted(ov, ov.u, True, True)
    assert_bits_equal(k1, k0, "dudt of the inferring call")
    assert_bits_equal(dt1, dt0, "dt")
    assert np.array_equal(t1, t0)
    st = ov.predictor_stats.cpu().numpy()
    assert st[0] == 0 and st[1] == int(t0.sum()) and st[2] == 0
    u1 = ov.u + float(dt0[0]) * 0.5 * _lib.torch().from_numpy(k0).to(ov.u.device)
    k2, dt2, t2 = _call_predicted(ov, u1, False, True)
    _call_predicted(seq, seq.u, True, False)
    k3, dt3, t3 = _call_predicted(seq, u1, False, False)
    assert_bits_equal(k2, k3, "dudt of the next call")
    assert_bits_equal(dt2, dt3, "dt of the next call")
    assert np.array_equal(t2, t3)
    pre = seq.predictor_stats.cpu().numpy()[0] // 2
    assert 0 < pre < n * n
    assert ov.predictor_stats.cpu().numpy()[0] == pre


def test_training_reduces_loss_on_solver_flags():
    """A few hundred Adam steps on flags of one quadrant run lower the loss."""
    torch = _lib.torch()
    n = 64
    u0 = ics.four_quadrants(n, seed=11)
    s = solver_from(u0, (n, n, 1), (3, 3, 0), (Ob, Ob, P), riemann_solver="hllc", a_posteriori_slope_limiting=True)
    s.integrator = "ssprk3"
    X, Y = [], []
    k, dt = _lib.empty(s.u.shape), _lib.empty((1,))
    for _ in range(6):
        for _ in range(5):
            s.take_step()
        s.stage(s.u, k, dt)
        X.append(s.predictor_features())
        Y.append(s._trouble[:, :, 0].clone())
    X, Y = torch.stack(X), torch.stack(Y)
    assert int(Y.sum().item()) > 0
    m = TroubledCellPredictor(seed=0)
    hist = train(m, X, Y, iters=300, batch=4, lr=3e-3, log_every=50)
    print(f"[train] loss {hist[0][1]:.4f} -> {hist[-1][1]:.4f}", flush=True)
    assert hist[-1][1] < 0.6 * hist[0][1]


# What the trained network gains on held-out four-quadrant problems (seeds 200-249 were never trained on):
# the check's flags left to the second pass, relative to the FP32 run without the predictor, summed over
# the seeds, stay below SECOND_PASS_SHARE_BOUND (the network moves fallbacks into the first pass); on a
# seed where it pre-flags almost nothing the share is one up to the few cells whose borderline tests
# differ between runs (SEED_SHARE_SLACK); the learned form's density differs from the plain FP64 run by
# less than LEARNED_L1_BOUND while its error against a 2x finer FP64 run stays within 1.5x of plain FP64's.
SECOND_PASS_SHARE_BOUND = 0.95
SEED_SHARE_SLACK = 1e-3
LEARNED_L1_BOUND = 1e-2
MATCHED_ERROR_RATIO = 1.5


def test_trained_predictor_gain_on_held_out_seeds():
    from fvhoe_amd.train_predictor import evaluate

    rows = evaluate(TroubledCellPredictor.load(), seeds=[201, 202, 205], n=128, t_end=0.05, log=lambda *a: None,
                    ref_seeds=3)
    done = [r for r in rows if "second_pass_share" in r]
    for r in rows:
        print("[gain] " + " ".join(f"{k}={v:.4g}" if isinstance(v, float) else f"{k}={v}" for k, v in r.items()),
              flush=True)
    assert len(done) >= 2, rows
    share = sum(r["second_pass"] for r in done) / sum(r["check_flags_fp32"] for r in done)
    print(f"[gain] second_pass_share over the seeds {share:.4f}", flush=True)
    assert share < SECOND_PASS_SHARE_BOUND
    for r in done:
        assert r["check_flags_fp32"] > 0
        assert r["second_pass_share"] < 1.0 + SEED_SHARE_SLACK, r
        assert r["l1_density_learned_fp32_vs_plain"] < LEARNED_L1_BOUND, r
        if "err_ratio_learned_fp32" in r:
            assert r["err_ratio_learned_fp32"] <= MATCHED_ERROR_RATIO, r
    assert sum("err_ratio_learned_fp32" in r for r in done) >= 1
