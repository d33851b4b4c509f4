"""RBC2D solvers on the GPU (drop-in for the reference's navier package)."""
from .rbc2d import NavierStokes
from .rbc2d_adj import NavierStokesAdjoint
from .rbc2d_base import NavierStokesBase, eval_Nu, eval_Nuvol

__all__ = ["NavierStokes", "NavierStokesAdjoint", "NavierStokesBase", "eval_Nu", "eval_Nuvol"]
