"""Energy batching, with the reference's semantics.

``get_batches`` follows ``src/qttools/utils/solvers_utils.py:7-32`` and
``get_section_sizes`` follows ``src/qttools/utils/mpi_utils.py:16-70``: the
stack of ``n`` energies is split into ``n // min(max_batch_size, n)`` balanced
sections, so individual batches can be larger than ``max_batch_size``
(e.g. ``n=10, max=3`` gives ``[4, 3, 3]``).
"""

from __future__ import annotations

import numpy as np


def get_section_sizes(
    num_elements: int, num_sections: int, strategy: str = "balanced"
) -> tuple[list, int]:
    """Sizes of ``num_sections`` sections covering ``num_elements`` elements.

    ``balanced`` puts the remainder one-by-one on the first sections;
    ``greedy`` fills sections to ``ceil(num_elements/num_sections)`` in order.
    Returns the section sizes and ``max(sizes) * num_sections``.
    """
    quotient, remainder = divmod(int(num_elements), int(num_sections))
    if strategy == "balanced":
        sizes = remainder * [quotient + 1] + (num_sections - remainder) * [quotient]
    elif strategy == "greedy":
        sizes = [0] * num_sections
        for i in range(num_sections):
            sizes[i] = min(quotient + min(remainder, 1), num_elements - sum(sizes))
    else:
        raise ValueError(f"Invalid strategy: {strategy}")
    return sizes, max(sizes) * num_sections


def get_batches(num_sections: int, max_batch_size: int) -> tuple[list, np.ndarray]:
    """Batch sizes and offsets for a stack of ``num_sections`` energies."""
    sizes, _ = get_section_sizes(
        num_elements=num_sections,
        num_sections=num_sections // min(max_batch_size, num_sections),
    )
    slices = np.hstack(([0], np.cumsum(np.asarray(sizes, dtype=np.int64))))
    return sizes, slices
