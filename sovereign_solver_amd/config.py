"""PDHG constants with the reference's values and names.

``solver/config.py:10,12,37-39,75``.  They are the defaults of the drop-in
calls, bound at import time exactly as the reference binds them
(``solver/lp/pdhg.py:49-55`` imports them by value and uses them as default
arguments).  ``PDHG_OVER_RELAXATION`` is kept for name compatibility; the
reference never reads it and neither does this engine.
"""

FEASIBILITY_TOL: float = 1e-6     # row-negation threshold of the reference form
PDHG_TOL: float = 1e-6            # relative KKT tolerance of solve_lp_pdhg
MAX_PDHG_ITERS: int = 100_000
PDHG_RESTART_PERIOD: int = 50     # KKT of the window average is evaluated every 50 iterations
PDHG_OVER_RELAXATION: float = 1.6  # unused, as in the reference
WARMSTART_FAILOVER_ITERS: int = 200  # a warm start no better than the cold start is dropped at once;
                                     # one whose KKT has not fallen, and is still at least a tenth of
                                     # the cold start's, by the first check at or after this many
                                     # iterations restarts cold (family and single engines)

# check cadence the reference derives from the restart period: max(10, min(P, 100))
CHECK_INTERVAL: int = max(10, min(PDHG_RESTART_PERIOD, 100))
