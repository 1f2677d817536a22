from quatrex_amd.greens_function_solver.rgf import RGF
from quatrex_amd.greens_function_solver.solver import GFSolver, OBCBlocks

__all__ = ["RGF", "GFSolver", "OBCBlocks"]
