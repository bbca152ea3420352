"""orctm_amd -- MI355X (gfx950) nonhydrostatic pressure Poisson solver for ORCTM.

The solver is the HIP library ``liborctm_nh.so`` behind the C ABI of ``include/orctm_nh.h``;
``orctm_amd.nh`` mirrors the reference's petsc_start / petsc_solve_3D / petsc_end interface on
top of it, ``orctm_amd.synth`` generates seeded inputs in the reference's Fortran layout.
"""
__all__ = ["abi", "synth", "nh"]
