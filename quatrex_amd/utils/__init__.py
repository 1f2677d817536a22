from quatrex_amd.utils.solvers_utils import get_batches, get_section_sizes

__all__ = ["get_batches", "get_section_sizes"]
