"""CPU oracle for the restarted-PDHG path.  TEST INFRASTRUCTURE ONLY.

Nothing in ``sovereign_solver_amd`` imports this package.  It is imported by
``tests/``, by ``__graft_entry__.smoke()`` (as the checker of the GPU result)
and by ``bench.py``'s CPU-baseline leg, and never serves as the product path.

Contents and what pins each part:

* ``pdhg_ref``  — FP64 NumPy/SciPy restatement of ``solver/lp/pdhg.py``
  (standard form, step sizes, KKT, iteration, averaging/restart, extraction).
  Pinned by the reference's published PDHG rows
  (``benchmarks/results/benchmark_comparison.csv:5,9,13,17,27,31,35``:
  status, objective to 8 digits and iteration count) on the reference's own
  instances, and by closed forms and invariants in ``tests/test_oracle.py``.
* ``yardstick`` — the reference KKT formula (``pdhg.py:184-196``) evaluated on
  the reference-form lift of a box-form primal/dual pair.  Pinned against
  ``pdhg_ref`` on explicitly lifted problems (the two must agree to rounding),
  and against hand-computed cases.
* ``simplex``   — FP64 bounded-variable revised simplex (the "own simplex"
  objective oracle).  Pinned against ``scipy.optimize.linprog(method="highs")``
  and against brute-force vertex enumeration on tiny LPs.
* ``mlp_ref``   — FP64 NumPy MLP forward/backward.  Pinned by central finite
  differences.
* ``farkas``    — the Farkas value of a dual vector (> 0 proves primal
  infeasibility), and the violations of a ray of decrease (dual
  infeasibility).  Pinned by hand-made certificates and rays, by F <= 0 for
  random duals of feasible LPs, and by the bound c'd >= -|y||viol(Kd)| -
  |r||viol_bounds(d)| for HiGHS's duals of bounded LPs.
* ``generators`` are not here: the workload generators are product code
  (``sovereign_solver_amd.workloads``) because users build the same families.
"""
