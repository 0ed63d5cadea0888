"""Shared test helpers: fixture loading and case -> oracle Scheme mapping."""

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def group(fix, prefix):
    return {k.split("/", 1)[1]: v for k, v in fix.items() if k.startswith(prefix + "/")}


def bcs(bc):
    return tuple(tuple(b) for b in bc)


def scheme_kwargs(n, p, bc, kw):
    """EulerSolver kwargs (fvhoe/solver.py:42-84) -> oracle.Scheme kwargs."""
    out = dict(n=tuple(n), p=tuple(p), bc=bcs(bc), gamma=kw.get("gamma", 1.4), CFL=kw.get("CFL", 0.8),
               riemann_solver=kw.get("riemann_solver", "hllc"),
               a_posteriori=kw.get("a_posteriori_slope_limiting", False),
               slope_limiter=kw.get("slope_limiter", "minmod"), force_trouble=kw.get("force_trouble", False),
               NAD=kw.get("NAD", 1e-2), NAD_mode=kw.get("NAD_mode", "global"),
               NAD_range=kw.get("NAD_range", "relative"), NAD_vars=kw.get("NAD_vars"), PAD=kw.get("PAD"),
               SED=kw.get("SED", True), convex=kw.get("convex", False))
    if "x" in kw:
        out["domain"] = (tuple(kw["x"]), (0, 1), (0, 1))
    return out


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))
