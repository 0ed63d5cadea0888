"""Learned troubled-cell predictor (no counterpart in FVHOE).

A three-layer valid 5x5 convolutional network, 4 -> 16 -> 16 -> 1 channels
(ReLU after the first two), maps the normalized primitive fv averages
(rho, vx, vy, P) of a stage input, extended by 6 ghost cells per side, to one
logit per cell: "the a-posteriori check will flag this cell".  In the learned
stage the cells it flags get MUSCL fluxes before the candidate is formed; the
plain NAD/PAD/SED check still runs on every cell and sends the cells it flags
to the fallback as well.

Everything runs in the HIP library: inference is one fused BF16-MFMA kernel
(FP32 accumulation), training uses FP32 convolution forward / backward, loss
and Adam kernels.  The network always sees its weights rounded to bfloat16.
"""

from __future__ import annotations

import json
import os
from typing import Optional

import numpy as np

from fvhoe_amd import _lib

K = 5
HALO = _lib.PRED_HALO
SHAPES = [("w1", (16, 4, K, K)), ("b1", (16,)), ("w2", (16, 16, K, K)), ("b2", (16,)),
          ("w3", (1, 16, K, K)), ("b3", (1,))]
OFFSETS = {}
_o = 0
for _name, _shape in SHAPES:
    OFFSETS[_name] = _o
    _o += int(np.prod(_shape))
NPARAMS = _o
assert NPARAMS == _lib.PRED_NPARAMS

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "quadrants_predictor.npz")


def he_init(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases (float32, NPARAMS)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("w"):
            parts.append(rng.standard_normal(shape).ravel() * np.sqrt(2.0 / (shape[1] * K * K)))
        else:
            parts.append(np.zeros(int(np.prod(shape))))
    return np.concatenate(parts).astype(np.float32)


def _at(t, offset_floats: int):
    return _lib.ctypes.c_void_p(t.data_ptr() + 4 * offset_floats)


class TroubledCellPredictor:
    """Network parameters on the device plus their packed inference form."""

    def __init__(self, theta=None, logit_threshold: float = 0.0, seed: int = 0, meta: dict = None):
        torch = _lib.torch()
        th = he_init(seed) if theta is None else np.asarray(theta, dtype=np.float32).ravel()
        if th.size != NPARAMS:
            raise ValueError(f"expected {NPARAMS} parameters, got {th.size}")
        self.theta = torch.from_numpy(th.copy()).to(_lib.device())
        self.packed = _lib.empty((_lib.PRED_PACKED_BYTES,), dtype=torch.uint8)
        self.logit_threshold = float(logit_threshold)
        self.meta = dict(meta or {})
        self.pack()

    # ------------------------------------------------------------ persistence
    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS) -> "TroubledCellPredictor":
        with np.load(path, allow_pickle=False) as z:
            theta = z["theta"]
            thr = float(z["logit_threshold"])
            meta = json.loads(str(z["meta"])) if "meta" in z else {}
        return cls(theta, thr, meta=meta)

    def save(self, path: str) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, theta=self.theta.detach().cpu().numpy(), logit_threshold=np.float64(self.logit_threshold),
                 meta=np.array(json.dumps(self.meta)))

    # -------------------------------------------------------------- inference
    def pack(self) -> None:
        """Refresh the packed BF16 inference weights from theta."""
        _lib.call("fvhoe_pred_pack", _lib.ptr(self.theta), _lib.ptr(self.packed), _lib.stream())

    def logits(self, X, n0: int, n1: int):
        """Logits (n0, n1) float32 of features X (4, n0+12, n1+12) (BF16-MFMA kernel)."""
        X = self._features_arg(X, n0, n1)
        out = _lib.empty((n0, n1), dtype=_lib.torch().float32)
        _lib.call("fvhoe_pred_infer", _lib.ptr(X), n0, n1, _lib.ptr(self.packed), float(self.logit_threshold),
                  None, _lib.ptr(out), _lib.stream())
        return out

    def flags(self, X, n0: int, n1: int):
        """Pre-flags (n0, n1) int8 = logit > logit_threshold."""
        X = self._features_arg(X, n0, n1)
        out = _lib.empty((n0, n1), dtype=_lib.torch().int8)
        _lib.call("fvhoe_pred_infer", _lib.ptr(X), n0, n1, _lib.ptr(self.packed), float(self.logit_threshold),
                  _lib.ptr(out), None, _lib.stream())
        return out

    @staticmethod
    def _features_arg(X, n0, n1):
        torch = _lib.torch()
        X, _ = _lib.as_device(X, torch.float32)
        if tuple(X.shape[-3:]) != (4, n0 + 2 * HALO, n1 + 2 * HALO):
            raise ValueError(f"features must have shape (4, {n0 + 12}, {n1 + 12}), got {tuple(X.shape)}")
        return X

    # --------------------------------------------------------------- training
    def forward_train(self, X):
        """FP32 training forward pass on a batch X (B, 4, H, W): logits (B, H-12, W-12)
        and the activations the backward pass needs."""
        torch = _lib.torch()
        X, _ = _lib.as_device(X, torch.float32)
        B, C, H, W = X.shape
        if C != 4 or H < 13 or W < 13:
            raise ValueError(f"bad feature batch shape {tuple(X.shape)}")
        a1 = _lib.empty((B, 16, H - 4, W - 4), dtype=torch.float32)
        a2 = _lib.empty((B, 16, H - 8, W - 8), dtype=torch.float32)
        z = _lib.empty((B, 1, H - 12, W - 12), dtype=torch.float32)
        th, s = self.theta, _lib.stream()
        _lib.call("fvhoe_pred_conv_forward", _lib.ptr(X), _at(th, OFFSETS["w1"]), _at(th, OFFSETS["b1"]),
                  _lib.ptr(a1), B, 4, 16, H, W, 1, s)
        _lib.call("fvhoe_pred_conv_forward", _lib.ptr(a1), _at(th, OFFSETS["w2"]), _at(th, OFFSETS["b2"]),
                  _lib.ptr(a2), B, 16, 16, H - 4, W - 4, 1, s)
        _lib.call("fvhoe_pred_conv_forward", _lib.ptr(a2), _at(th, OFFSETS["w3"]), _at(th, OFFSETS["b3"]),
                  _lib.ptr(z), B, 16, 1, H - 8, W - 8, 0, s)
        return z[:, 0], (X, a1, a2)

    def loss_and_dlogits(self, z, labels, pos_weight: float):
        """Mean class-weighted BCE of logits z against int8 labels, and dL/dz."""
        torch = _lib.torch()
        labels, _ = _lib.as_device(labels, torch.int8)
        if labels.shape != z.shape:
            raise ValueError(f"labels {tuple(labels.shape)} do not match logits {tuple(z.shape)}")
        dz = torch.empty_like(z)
        loss = _lib.zeros((1,), dtype=torch.float32)
        _lib.call("fvhoe_pred_loss", _lib.ptr(z), _lib.ptr(labels), float(pos_weight), z.numel(), _lib.ptr(dz),
                  _lib.ptr(loss), _lib.stream())
        return loss, dz

    def backward(self, dz, cache):
        """Gradient (NPARAMS float32) of the loss whose dL/dz is dz."""
        torch = _lib.torch()
        X, a1, a2 = cache
        B, _, H, W = X.shape
        g = _lib.zeros((NPARAMS,), dtype=torch.float32)
        da2 = torch.empty_like(a2)
        da1 = torch.empty_like(a1)
        th, s = self.theta, _lib.stream()
        dz4 = dz.reshape(B, 1, H - 12, W - 12).contiguous()
        _lib.call("fvhoe_pred_conv_backward", _lib.ptr(a2), _lib.ptr(dz4), _at(th, OFFSETS["w3"]), _lib.ptr(da2),
                  _at(g, OFFSETS["w3"]), _at(g, OFFSETS["b3"]), B, 16, 1, H - 8, W - 8, s)
        _lib.call("fvhoe_pred_conv_backward", _lib.ptr(a1), _lib.ptr(da2), _at(th, OFFSETS["w2"]), _lib.ptr(da1),
                  _at(g, OFFSETS["w2"]), _at(g, OFFSETS["b2"]), B, 16, 16, H - 4, W - 4, s)
        _lib.call("fvhoe_pred_conv_backward", _lib.ptr(X), _lib.ptr(da1), _at(th, OFFSETS["w1"]), None,
                  _at(g, OFFSETS["w1"]), _at(g, OFFSETS["b1"]), B, 4, 16, H, W, s)
        return g


class Adam:
    """Adam on the predictor's FP32 master weights (one HIP kernel per step)."""

    def __init__(self, model: TroubledCellPredictor, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8):
        self.model = model
        self.lr, self.b1, self.b2, self.eps = float(lr), float(betas[0]), float(betas[1]), float(eps)
        self.m = _lib.zeros((NPARAMS,), dtype=_lib.torch().float32)
        self.v = _lib.zeros((NPARAMS,), dtype=_lib.torch().float32)
        self.t = 0

    def step(self, grad) -> None:
        self.t += 1
        _lib.call("fvhoe_pred_adam", _lib.ptr(self.model.theta), _lib.ptr(grad), _lib.ptr(self.m), _lib.ptr(self.v),
                  NPARAMS, self.lr, self.b1, self.b2, self.eps, self.t, _lib.stream())


def train(model: TroubledCellPredictor, X, Y, iters: int = 2000, batch: int = 8, lr: float = 2e-3,
          pos_weight: Optional[float] = None, seed: int = 0, log_every: int = 0, lr_decay_at=(0.6, 0.85),
          max_pos_weight: float = 10.0, log=None):
    """Train on features X (N, 4, H, W) float32 and labels Y (N, H-12, W-12) int8,
    both device tensors; random mini-batches, Adam, step decay of the rate.
    pos_weight defaults to half the negative/positive ratio, capped at
    max_pos_weight.  Returns the list of (iteration, loss)."""
    torch = _lib.torch()
    N = X.shape[0]
    if pos_weight is None:
        frac = float(Y.float().mean().item())
        pos_weight = float(min(max_pos_weight, max(1.0, 0.5 * (1 - frac) / max(frac, 1e-6))))
    opt = Adam(model, lr=lr)
    gen = torch.Generator(device="cpu")
    gen.manual_seed(seed)
    hist = []
    for it in range(iters):
        if it in (int(lr_decay_at[0] * iters), int(lr_decay_at[1] * iters)):
            opt.lr *= 0.3
        idx = torch.randint(0, N, (batch,), generator=gen).to(X.device)
        xb = X.index_select(0, idx).contiguous()
        yb = Y.index_select(0, idx).contiguous()
        z, cache = model.forward_train(xb)
        loss, dz = model.loss_and_dlogits(z, yb, pos_weight)
        g = model.backward(dz, cache)
        opt.step(g)
        if log_every and (it % log_every == 0 or it == iters - 1):
            hist.append((it, float(loss.item())))
            if log is not None:
                log(f"# iter {it} loss {hist[-1][1]:.5f}")
    model.pack()
    model.meta.update({"pos_weight": pos_weight, "iters": iters, "batch": batch, "lr": lr})
    return hist


def confusion(pred, label) -> dict:
    """Counts of a pre-flag map against the check's flags (device int8 tensors)."""
    p = pred.bool()
    y = label.bool()
    tp = int((p & y).sum().item())
    fp = int((p & ~y).sum().item())
    fn = int((~p & y).sum().item())
    return {"tp": tp, "fp": fp, "fn": fn, "recall": tp / max(tp + fn, 1), "precision": tp / max(tp + fp, 1)}
