"""Mixed-integer consumers of the batched PDHG engine."""

from .branch_and_bound import MILPSolveResult, solve_milp_batched

__all__ = ["MILPSolveResult", "solve_milp_batched"]
