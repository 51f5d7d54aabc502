"""MI355X-native restarted-PDHG LP engine.

Drop-in for the PDHG entry points of the reference solver package
(``solver/__init__.py:22-38`` exports ``solve_lp_pdhg``,
``solve_lp_pdhg_gpu``, ``get_available_device``, ``Problem``, ``read_mps``):
the same names and call signatures, computed by hand-written HIP kernels for
gfx950 behind the C ABI of ``libpdlp_hip.so``.
"""

from .problem import Problem
from .result import SolveResult
from .io.mps import read_mps
from .lp.pdhg import solve_lp_pdhg
from .lp.pdhg_gpu import get_available_device, solve_lp_pdhg_gpu
from .lp.batch import solve_lp_pdhg_batch
from .lp.branching import branch_strong_batched
from .milp import MILPSolveResult, solve_milp_batched

__all__ = ["Problem", "SolveResult", "read_mps", "solve_lp_pdhg", "solve_lp_pdhg_gpu", "get_available_device",
           "solve_lp_pdhg_batch", "branch_strong_batched", "MILPSolveResult", "solve_milp_batched"]
