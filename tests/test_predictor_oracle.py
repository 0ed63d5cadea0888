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
