"""NumPy reference of the troubled-cell predictor (test infrastructure only).

The predictor has no counterpart in FVHOE.  Network: three valid 5x5
convolutions 4 -> 16 -> 16 -> 1 channels with ReLU after the first two, on
feature maps padded by 6 cells; weights are used rounded to bfloat16 (round
to nearest even), activations and accumulation in float64 here.  Loss:
class-weighted binary cross-entropy on logits, mean over pixels.

Pinned by finite differences of the loss (tests/test_predictor_oracle.py).
"""

from __future__ import annotations

import numpy as np

K = 5
CH = (4, 16, 16, 1)
HALO = 6


def param_shapes():
    return [("w1", (16, 4, K, K)), ("b1", (16,)), ("w2", (16, 16, K, K)), ("b2", (16,)),
            ("w3", (1, 16, K, K)), ("b3", (1,))]


def n_params() -> int:
    return sum(int(np.prod(s)) for _, s in param_shapes())


def unpack(theta: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in param_shapes():
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even) -> float32."""
    b = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (b >> 16) & 1
    r = ((b + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def conv_valid(x: np.ndarray, w: np.ndarray, b: np.ndarray) -> np.ndarray:
    """x (B, Cin, H, W), w (Cout, Cin, K, K) -> (B, Cout, H-K+1, W-K+1), cross-correlation."""
    B, C, H, W = x.shape
    Ho, Wo = H - K + 1, W - K + 1
    out = np.zeros((B, w.shape[0], Ho, Wo))
    for dy in range(K):
        for dx in range(K):
            patch = x[:, :, dy:dy + Ho, dx:dx + Wo]
            out += np.einsum("bchw,oc->bohw", patch, w[:, :, dy, dx])
    return out + b[None, :, None, None]


def forward(theta: np.ndarray, x: np.ndarray, rounding: bool = True, act_bf16: bool = False):
    """Logits (B, H-12, W-12).  rounding=False uses the float weights as they are
    (for finite-difference checks of backward).  act_bf16 also rounds the input
    and the two hidden activations to bfloat16, as the inference kernel stores
    them (its BF16 MFMA data path)."""
    p = {k: (bf16_round(v) if (rounding and k.startswith("w")) else v).astype(np.float64)
         for k, v in unpack(theta).items()}
    act = (lambda a: bf16_round(a.astype(np.float32)).astype(np.float64)) if act_bf16 else (lambda a: a)
    x = act(x)
    z1 = conv_valid(x, p["w1"], p["b1"])
    a1 = act(np.maximum(z1, 0))
    z2 = conv_valid(a1, p["w2"], p["b2"])
    a2 = act(np.maximum(z2, 0))
    z3 = conv_valid(a2, p["w3"], p["b3"])[:, 0]
    return z3, (x, z1, a1, z2, a2, p)


def loss(z: np.ndarray, y: np.ndarray, pos_weight: float) -> float:
    sp = np.logaddexp(0, -z)  # softplus(-z)
    l = pos_weight * y * sp + (1 - y) * (z + sp)
    return float(l.mean())


def backward(z, cache, y, pos_weight):
    x, z1, a1, z2, a2, p = cache
    n = z.size
    s = 1 / (1 + np.exp(-z))
    dz = (pos_weight * y * (s - 1) + (1 - y) * s) / n
    g = {}

    def conv_grads(inp, dout, w):
        B, C, H, W = inp.shape
        Ho, Wo = dout.shape[2], dout.shape[3]
        dw = np.zeros_like(w)
        din = np.zeros_like(inp)
        for dy in range(K):
            for dx in range(K):
                patch = inp[:, :, dy:dy + Ho, dx:dx + Wo]
                dw[:, :, dy, dx] = np.einsum("bohw,bchw->oc", dout, patch)
                din[:, :, dy:dy + Ho, dx:dx + Wo] += np.einsum("bohw,oc->bchw", dout, w[:, :, dy, dx])
        return dw, dout.sum(axis=(0, 2, 3)), din

    d3 = dz[:, None]
    g["w3"], g["b3"], da2 = conv_grads(a2, d3, p["w3"])
    d2 = da2 * (z2 > 0)
    g["w2"], g["b2"], da1 = conv_grads(a1, d2, p["w2"])
    d1 = da1 * (z1 > 0)
    g["w1"], g["b1"], _ = conv_grads(x, d1, p["w1"])
    return np.concatenate([g[name].ravel() for name, _ in param_shapes()])


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in param_shapes():
        if name.startswith("w"):
            fan_in = shape[1] * K * K
            parts.append(rng.standard_normal(shape).ravel() * np.sqrt(2.0 / fan_in))
        else:
            parts.append(np.zeros(int(np.prod(shape))))
    return np.concatenate(parts).astype(np.float32)


def features(s, u: np.ndarray) -> np.ndarray:
    """Predictor input for the 2-D stage input u: (4, n0+12, n1+12) float32.

    The primitive fv averages of u with 3 ghost cells (fvhoe/solver.py:753-792
    with gw=3, as in revise_fluxes, fvhoe/solver.py:681-687) give the global
    (min, max) per variable (fvhoe/slope_limiting.py:168-170); the interior
    averages of (rho, vx, vy, P), extended by 6 cells per side with the
    scheme's boundary conditions, are scaled to (w - min) * (1 / (max - min))
    (0 where max = min), in float64, then rounded to float32."""
    from oracle import fvhoe_oracle as O

    w3 = O.interpolate_primitives_from_conservatives(s, u, s.p, (3, 3, 0))
    lo = w3.min(axis=(1, 2, 3))
    hi = w3.max(axis=(1, 2, 3))
    we = O.apply_bc(w3[:, 3:-3, 3:-3, :], (HALO, HALO, 0), s.bc)[..., 0]
    out = []
    for v in (0, 1, 2, 4):
        rng = hi[v] - lo[v]
        out.append((we[v] - lo[v]) * (1.0 / rng) if rng > 0 else np.zeros_like(we[v]))
    return np.stack(out).astype(np.float32)


# ---------------------------------------------------------------------------
# The network on planar states (1-D runs of the ensemble).  A 1-D run along x
# seen as a 2-D state that does not vary along y gives the network input that
# does not vary along the image's second axis, and each valid 5x5
# cross-correlation of such an input equals a 1-D one along the first axis
# with the kernel summed over the second axis (dx); so does every later layer.
# The planar network is that 1-D network, with the summed kernels of layers 1
# and 2 rounded to bfloat16 once more (the inference kernel's MFMA operands)
# and layer 3's kept in float32; activations rounded to bfloat16 between
# layers, as in the 2-D kernel.
# ---------------------------------------------------------------------------


def planar_weights(theta: np.ndarray) -> dict:
    """w1p (16, 4, 5), w2p (16, 16, 5): bf16(sum over dx of bf16(w)) with the sum in float32 in dx order;
    w3p (16, 5): the float32 sum; biases as they are."""
    p = unpack(np.asarray(theta, dtype=np.float32))
    out = {}
    for name in ("w1", "w2", "w3"):
        w = bf16_round(p[name])
        acc = np.zeros(w.shape[:3], dtype=np.float32)
        for dx in range(K):
            acc = (acc + w[..., dx]).astype(np.float32)
        out[name + "p"] = acc if name == "w3" else bf16_round(acc)
    out["w3p"] = out["w3p"][0]
    for name in ("b1", "b2", "b3"):
        out[name] = p[name].astype(np.float32)
    return out


def conv1d_valid(x: np.ndarray, w: np.ndarray, b: np.ndarray) -> np.ndarray:
    """x (Cin, L), w (Cout, Cin, K) -> (Cout, L-K+1), cross-correlation, float64."""
    L = x.shape[1] - K + 1
    out = np.zeros((w.shape[0], L))
    for t in range(K):
        out += np.einsum("oc,cl->ol", w[:, :, t].astype(np.float64), x[:, t:t + L])
    return out + b.astype(np.float64)[:, None]


def forward_planar(theta: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Logits (L-12,) of the planar network on features x (4, L) (float64 arithmetic, input and hidden
    activations rounded to bfloat16 as the kernel stores them)."""
    q = planar_weights(theta)
    act = lambda a: bf16_round(a.astype(np.float32)).astype(np.float64)  # noqa: E731
    a0 = act(np.asarray(x, dtype=np.float64))
    a1 = act(np.maximum(conv1d_valid(a0, q["w1p"], q["b1"]), 0))
    a2 = act(np.maximum(conv1d_valid(a1, q["w2p"], q["b2"]), 0))
    return conv1d_valid(a2, q["w3p"][None], q["b3"])[0]


def features_1d(w: np.ndarray, rng_lo: np.ndarray, rng_hi: np.ndarray, bc) -> np.ndarray:
    """Planar-network input of one 1-D run: w (5, n) primitive fv averages of the interior cells, the
    run's NAD range (min, max per variable over its cells and 3 ghosts on each side), bc = (lo, hi)
    boundary kinds.  (rho, vx, vy, P) at cells -6 .. n+5, ghosts mapped into the interior as the 2-D
    features map them (periodic wrap, free = nearest cell, otherwise mirror, vx negated across a
    reflective side), scaled to (w - min) * (1 / (max - min)) in float64 (0 where max = min), rounded to
    float32: (4, n + 12)."""
    n = w.shape[1]
    c = np.arange(-HALO, n + HALO)
    src = c.copy()
    flip = np.zeros(c.shape, dtype=bool)
    for side, sel, edge in ((bc[0], c < 0, 0), (bc[1], c >= n, n - 1)):
        if side == "periodic":
            src[sel] = c[sel] % n
        elif side == "free":
            src[sel] = edge
        else:
            src[sel] = np.where(c[sel] < 0, -1 - c[sel], 2 * n - 1 - c[sel])
            flip[sel] = side == "reflective"
    out = []
    for v in (0, 1, 2, 4):
        x = w[v][src].astype(np.float64)
        if v == 1:
            x = np.where(flip, -x, x)
        r = rng_hi[v] - rng_lo[v]
        out.append((x - rng_lo[v]) * (1.0 / r) if r > 0 else np.zeros_like(x))
    return np.array(out).astype(np.float32)
