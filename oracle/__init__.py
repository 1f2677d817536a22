"""CPU oracle for the RGF selected inversion / selected solve.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or, for ``cpu_baseline``, as the reference CPU path being timed).
The product path (``quatrex_amd``) never imports it.

Contents:

* :mod:`oracle.rgf_ref` -- numpy restatement of ``RGF.selected_inv`` and
  ``RGF.selected_solve`` (``src/qttools/greens_function_solver/rgf.py``) with
  the reference's operation order, on a minimal block-tridiagonal container
  that mimics ``DSDBSparse.blocks`` / ``.stack``;
* :mod:`oracle.dense` -- dense oracles: ``numpy.linalg.inv`` of the full
  matrix (selected by the tridiagonal pattern), ``G^R Sigma G^R^H`` for the
  lesser/greater functions, and a brute-force Meir-Wingreen current from dense
  left-connected inverses;
* :mod:`oracle.batching` -- ``get_batches`` / ``get_section_sizes``.

How each function is pinned (by the tests under ``tests/``):

* ``selected_inv``, ``selected_solve`` (G^R all tridiagonal blocks, G^<, G^>):
  against the dense oracle at relative Frobenius error <= 1e-10 per block, on
  the reference test axes (constant and mixed block sizes, batch sizes 1/2/100,
  1-D/2-D/3-D stacks, out/None, return_retarded) -- the reference's own test
  method (``tests/qttools/greens_function_solver/test_selected_solve.py:56-64``);
* anti-Hermitian projection and ``G^<_ji = -G^<_ij^H``: dense equality plus the
  invariant that exact G^<, G^> are anti-Hermitian;
* current: brute force from dense left-connected inverses on tiny inputs, and
  current conservation in the ballistic case;
* batching: the closed-form examples of the reference docstring.

The reference itself is not importable here (Python >= 3.13, mpi4py, numba);
no reference code is run or copied.
"""
