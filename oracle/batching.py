"""Oracle restatement of the energy batching (test infrastructure only).

Follows ``src/qttools/utils/mpi_utils.py:16-70`` (``get_section_sizes``) and
``src/qttools/utils/solvers_utils.py:7-32`` (``get_batches``).  Pinned by the
docstring examples of ``get_section_sizes`` (``(10, 3) -> [4, 3, 3]`` balanced,
``[4, 4, 2]`` greedy) and by brute-force properties in ``tests/``.
"""

import numpy as np


def section_sizes(n: int, k: int, strategy: str = "balanced"):
    q, r = divmod(n, k)
    if strategy == "balanced":
        out = [q + 1 if s < r else q for s in range(k)]
    elif strategy == "greedy":
        out = []
        left = n
        for _ in range(k):
            take = min(q + (1 if r else 0), left)
            out.append(take)
            left -= take
    else:
        raise ValueError(strategy)
    return out, max(out) * k


def batches(n: int, max_batch_size: int):
    k = n // min(max_batch_size, n)
    sizes, _ = section_sizes(n, k)
    return sizes, np.concatenate(([0], np.cumsum(sizes))).astype(np.int64)
