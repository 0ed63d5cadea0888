"""CPU oracle for the per-RK-stage update of FVHOE (test infrastructure only).

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.  It restates, op for op and in the same floating-point order, the
reference's NumPy path (citations are ``file:line`` into jpalafou/FVHOE).

Pinning (what checks each function besides the restatement itself):

* stencil weights          -- exact-rational polynomial exactness (tests/test_oracle_closed_forms.py),
                              C table == oracle bit for bit (tests/test_abi_cpu.py)
* stencil sweeps, centers, averages -- brute force, polynomial centres/averages (test_oracle_closed_forms)
                              + golden stages (tests/test_oracle_golden.py)
* primitives/conservatives -- closed form and round trip + golden stages
* sound speed / hydro_dt   -- closed form + golden stages
* llf                      -- Toro's textbook form, consistency, mirror symmetry + golden stages
* hllc                     -- consistency, supersonic limits, mirror symmetry + golden stages
* MUSCL/minmod/moncen      -- brute force on tiny inputs + golden force_trouble stages
* detect_troubled_cells, SED, broadcast -- golden vectors written by tests/golden/make_golden.py
* boundary conditions      -- brute-force index maps (test_oracle_closed_forms), the reference's
                              Dirichlet / "ic" cases (tests/test_oracle_boundaries.py) + golden stages
* whole stage / SSP-RK3 trajectories -- golden vectors written by tests/golden/make_golden.py
* FP32 bulk form (no reference counterpart) -- scalar float32 brute force, float64 recovery rounded once,
                              float64 stages to float32 rounding (tests/test_oracle_fp32.py)
"""

from oracle.fvhoe_oracle import *  # noqa: F401,F403
