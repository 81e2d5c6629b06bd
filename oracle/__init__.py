"""CPU oracle for the Monte Carlo path-simulation hot path.

TEST INFRASTRUCTURE ONLY. This package is a plain numpy / Python-int
restatement of the tf-quant-finance algorithms on the simulation path
(Sobol sampling, time grids, Euler-Maruyama, Heston QE, exact GBM,
Longstaff-Schwartz, closed-form prices).  Each function cites the reference
file:line it follows.  It exists to check the HIP path, and may be imported
only by `tests/`, `__graft_entry__.smoke()` and the CPU-baseline leg of
`bench.py`.  Nothing in the `tqf` package imports it and no product code path
routes through it.

Parity pinning (how each function is checked against something other than a
retyped copy of itself) is recorded in each module's header.  Functions with
no independent check say "parity unpinned".
"""
