"""Open boundary conditions on the device (Sancho-Rubio surface Green's
function, memoising boundary system, contact self-energies)."""

from quatrex_amd.boundary_conditions.obc import (
    OBCSolver,
    OBCSystem,
    SanchoRubio,
    compute_contact_obc,
    compute_obc,
)

__all__ = ["OBCSolver", "OBCSystem", "SanchoRubio", "compute_contact_obc", "compute_obc"]
