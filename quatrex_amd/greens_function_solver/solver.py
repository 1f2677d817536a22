"""GF-solver plugin interface, as in ``src/qttools/greens_function_solver/solver.py``.

``OBCBlocks`` (``solver.py:12-28``) holds per-block lists of optional open
boundary blocks; ``GFSolver`` (``solver.py:31-113``) is the abstract plugin with
``selected_inv`` and ``selected_solve``.  Same names, signatures and return
conventions, so a caller can swap the solver class without other changes.
"""

from __future__ import annotations

from abc import ABC, abstractmethod


class OBCBlocks:
    """Open-boundary blocks: ``retarded``, ``lesser``, ``greater`` lists of length
    ``num_blocks`` whose entries are ``None`` or stacked ``(*stack, b_i, b_i)``
    arrays indexed by the energy slice."""

    def __init__(self, num_blocks: int):
        self.retarded: list = [None] * num_blocks
        self.lesser: list = [None] * num_blocks
        self.greater: list = [None] * num_blocks


class GFSolver(ABC):
    """Abstract Green's-function solver."""

    @abstractmethod
    def selected_inv(self, a, obc_blocks: OBCBlocks | None = None, out=None):
        """Selected inversion of a block-tridiagonal matrix; returns the new
        matrix when ``out`` is None, else None."""

    @abstractmethod
    def selected_solve(self, a, sigma_lesser, sigma_greater, obc_blocks: OBCBlocks | None = None,
                       out=None, return_retarded: bool = False, return_current: bool = False):
        r"""Selected elements of ``X^<> = A^{-1} Sigma^<> A^{-H}`` (and ``A^{-1}``)."""
