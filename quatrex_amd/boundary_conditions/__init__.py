"""Open boundary conditions on the device (Sancho-Rubio and spectral surface
Green's functions, memoising boundary system, contact self-energies)."""

from quatrex_amd.boundary_conditions.obc import (
    OBCSolver,
    OBCSystem,
    SanchoRubio,
    compute_contact_obc,
    compute_obc,
)
from quatrex_amd.boundary_conditions.spectral import Spectral

__all__ = ["OBCSolver", "OBCSystem", "SanchoRubio", "Spectral", "compute_contact_obc", "compute_obc"]
