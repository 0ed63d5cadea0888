"""The drop-in boundary by name: every public function and class method of
the reference modules on the per-stage path (recorded from the reference by
tests/golden/make_golden.py into signatures.json) has a counterpart in
fvhoe_amd with the same parameter names, kinds, order and defaults; extra
parameters of this build come after the reference's, with defaults."""

import importlib
import inspect
import json
import os
import re

import pytest

from helpers import GOLDEN

with open(os.path.join(GOLDEN, "signatures.json")) as f:
    SIGS = json.load(f)

# reference names with no counterpart, and why
NOT_PROVIDED = {
    ("solver", "EulerSolver", "plot_fields"): "matplotlib visualisation, off the per-stage path",
}


def _norm_default(d):
    """Object defaults compare by class name (fvhoe.initial_conditions.Square vs fvhoe_amd's)."""
    if d is None:
        return None
    m = re.match(r"<([\w.]+)\.(\w+) object>", re.sub(r" at 0x[0-9a-fA-F]+", "", d))
    return f"<{m.group(2)}>" if m else d


def _check(ref_params, obj, what):
    sig = inspect.signature(obj)
    ours = list(sig.parameters.values())
    assert len(ours) >= len(ref_params), (what, [p.name for p in ours])
    for (name, kind, default), p in zip(ref_params, ours):
        assert p.name == name, (what, name, p.name)
        assert str(p.kind) == kind, (what, name, kind, str(p.kind))
        d = None if p.default is inspect.Parameter.empty else repr(p.default)
        assert _norm_default(d) == _norm_default(default), (what, name, default, d)
    for p in ours[len(ref_params):]:  # this build's additions: optional, after the reference's
        assert p.default is not inspect.Parameter.empty or p.kind in (p.VAR_POSITIONAL, p.VAR_KEYWORD), (what, p.name)


CASES = [(m, name) for m in sorted(SIGS) for name in sorted(SIGS[m])]


@pytest.mark.parametrize("module,name", CASES, ids=[f"{m}.{n}" for m, n in CASES])
def test_counterpart_has_the_reference_signature(module, name):
    mod = importlib.import_module(f"fvhoe_amd.{module}")
    entry = SIGS[module][name]
    assert hasattr(mod, name), f"fvhoe_amd.{module} has no {name}"
    obj = getattr(mod, name)
    if entry["kind"] == "function":
        assert callable(obj) and not inspect.isclass(obj), name
        _check(entry["params"], obj, f"{module}.{name}")
        return
    assert inspect.isclass(obj), name
    for meth, params in entry["methods"].items():
        if (module, name, meth) in NOT_PROVIDED:
            continue
        assert hasattr(obj, meth), f"{module}.{name}.{meth}"
        _check(params, getattr(obj, meth), f"{module}.{name}.{meth}")


def test_every_reference_module_is_recorded():
    assert set(SIGS) == {"fv", "stencils", "hydro", "riemann_solvers", "solver", "boundary_conditions",
                         "slope_limiting", "smooth_extrema_detection", "ode"}
    assert sum(len(v) for v in SIGS.values()) >= 36
