"""Pins the predictor's NumPy oracle: backward against central finite
differences of the loss, bf16 rounding against exact cases."""

import numpy as np

from oracle import predictor_oracle as PO


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -3.140625, 65504.0, 1.0 + 3 * 2 ** -8],
                 dtype=np.float32)
    r = PO.bf16_round(x)
    # 1+2^-8 is a tie -> even (1.0); 1+2^-7+2^-9 rounds down to 1+2^-7; 1+3*2^-8 tie -> up to 1+2^-6
    assert list(r) == [1.0, 1.0, 1.0 + 2 ** -7, -3.140625, 65536.0, 1.0 + 2 ** -6]


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    theta = PO.init_params(1).astype(np.float64)
    theta[-1] = -0.5
    x = rng.standard_normal((2, 4, 15, 16))
    y = (rng.random((2, 3, 4)) < 0.3).astype(np.float64)
    z, cache = PO.forward(theta, x, rounding=False)
    g = PO.backward(z, cache, y, pos_weight=3.0)
    idx = rng.choice(theta.size, 40, replace=False)
    idx = np.concatenate([idx, [theta.size - 1, 1600, 1600 + 16 + 7]])
    for i in idx:
        h = 1e-6
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fp = PO.loss(PO.forward(tp, x, rounding=False)[0], y, 3.0)
        fm = PO.loss(PO.forward(tm, x, rounding=False)[0], y, 3.0)
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_conv_valid_brute_force():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1, 2, 7, 6))
    w = rng.standard_normal((3, 2, 5, 5))
    b = rng.standard_normal(3)
    out = PO.conv_valid(x, w, b)
    for o in range(3):
        for i in range(3):
            for j in range(2):
                ref = b[o] + sum(w[o, c, dy, dx] * x[0, c, i + dy, j + dx]
                                 for c in range(2) for dy in range(5) for dx in range(5))
                assert abs(out[0, o, i, j] - ref) < 1e-12


def test_features_invariants():
    """The predictor features: the range-defining cells map to exactly 0 and to
    1 within rounding; ghost cells follow the boundary maps (periodic wrap,
    mirror, mirror with the normal velocity negated)."""
    from oracle import fvhoe_oracle as O
    from fvhoe_amd import initial_conditions as ics

    n = 24
    u = ics.four_quadrants(n, seed=3)
    for kind in ("periodic", "outflow", "reflective"):
        s = O.Scheme(n=(n, n, 1), p=(3, 3, 0), gamma=1.4, a_posteriori=True,
                     bc=((kind, kind), (kind, kind), ("periodic", "periodic")))
        X = PO.features(s, u)
        assert X.shape == (4, n + 12, n + 12) and X.dtype == np.float32
        w3 = O.interpolate_primitives_from_conservatives(s, u, s.p, (3, 3, 0))[..., 0]
        core = X[:, 3:-3, 3:-3]  # the interior and 3 ghost cells per side
        for k, v in enumerate((0, 1, 2, 4)):
            if w3[v].max() > w3[v].min():
                assert core[k].min() >= 0.0 and core[k].max() <= 1.0 + 1e-6
        inner = X[:, 6:-6, 6:-6]
        if kind == "periodic":
            assert np.array_equal(X[:, :6, 6:-6], inner[:, -6:, :])
        else:
            mirrored = X[:, :6, 6:-6][:, ::-1, :]
            same = (0, 2, 3) if kind == "reflective" else (0, 1, 2, 3)
            assert np.array_equal(mirrored[list(same)], inner[list(same), :6, :])
            if kind == "reflective":  # vx negated about the wall: f(-v) = 2 f(0) - f(v) in exact arithmetic
                w = O.interpolate_primitives_from_conservatives(s, u, s.p, (0, 0, 0))[..., 0]
                lo, hi = w3[1].min(), w3[1].max()
                ref = ((-w[1, :6, :] - lo) * (1.0 / (hi - lo))).astype(np.float32)
                assert np.array_equal(mirrored[1], ref)


def test_planar_network_is_the_2d_network_on_y_invariant_inputs():
    """The planar network (kernels summed over the second image axis) against the 2-D oracle network on
    inputs constant along that axis: equal up to the second bfloat16 rounding of the summed kernels (a
    relative weight change <= 2^-9); summing over the wrong axis, or taking one column of the kernel,
    misses by far more."""
    rng = np.random.default_rng(0)
    for seed in (1, 3, 7):
        th = PO.init_params(seed)
        th[-1] = 0.1
        x = rng.random((4, 44)).astype(np.float32)
        z1 = PO.forward_planar(th, x)
        X = np.repeat(x[None, :, :, None], 13, axis=3).astype(np.float64)
        z2 = PO.forward(th, X, rounding=True, act_bf16=True)[0][0]
        assert z2.shape == (32, 1) and z1.shape == (32,)
        scale = np.abs(z2).max()
        err = np.abs(z2[:, 0] - z1).max()
        assert err <= 2e-2 * scale, (seed, err, scale)
        # the same network with the kernels summed over the first axis instead
        p = PO.unpack(th.copy())
        for name in ("w1", "w2", "w3"):
            p[name][...] = np.swapaxes(p[name], 2, 3)
        th_t = np.concatenate([p[k].ravel() for k, _ in PO.param_shapes()]).astype(np.float32)
        wrong = np.abs(z2[:, 0] - PO.forward_planar(th_t, x)).max()
        assert wrong > 10 * err, (seed, wrong, err)


def test_planar_weights_closed_form():
    th = PO.init_params(5)
    q = PO.planar_weights(th)
    p = PO.unpack(th)
    w1 = PO.bf16_round(p["w1"]).astype(np.float64).sum(axis=3)
    assert q["w1p"].shape == (16, 4, 5) and q["w2p"].shape == (16, 16, 5) and q["w3p"].shape == (16, 5)
    assert np.abs(q["w1p"] - w1).max() <= 2 ** -8 * np.abs(w1).max()
    # bfloat16 values: the low 16 bits are zero
    assert not np.any(q["w2p"].view(np.uint32) & 0xFFFF)


def test_features_1d_boundary_maps():
    n = 10
    w = np.arange(5 * n, dtype=np.float64).reshape(5, n) + 1.0
    lo, hi = w.min(axis=1), w.max(axis=1)
    for bc, left, right in ((("periodic", "periodic"), [n - 6 + k for k in range(6)], list(range(6))),
                            (("free", "free"), [0] * 6, [n - 1] * 6),
                            (("outflow", "reflective"), [5 - k for k in range(6)], [n - 1 - k for k in range(6)])):
        X = PO.features_1d(w, lo, hi, bc)
        assert X.shape == (4, n + 12)
        want = lambda src: ((w[0][src] - lo[0]) / (hi[0] - lo[0])).astype(np.float32)  # noqa: E731
        assert np.array_equal(X[0, :6], want(np.array(left))), bc
        assert np.array_equal(X[0, -6:], want(np.array(right))), bc
        assert np.array_equal(X[0, 6:-6], want(np.arange(n)))
    X = PO.features_1d(w, lo, hi, ("reflective", "reflective"))
    vx = (-w[1][[5 - k for k in range(6)]] - lo[1]) / (hi[1] - lo[1])
    assert np.allclose(X[1, :6], vx)
    w[2] = 0.0
    assert not np.any(PO.features_1d(w, w.min(axis=1), w.max(axis=1), ("periodic", "periodic"))[2])
