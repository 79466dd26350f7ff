"""MI355X (gfx950) implementation of the linear-solve hot path of PARALLEL_FEM_CODES.

The product is ``libfemgpu.so`` (HIP kernels + C ABI, ``include/femgpu.h``); ``fem`` mirrors the
reference interface on top of it for Python callers and tests.
"""
from . import native
from .fem import (CsrMatrix, Device, EbeMatrix, Fill_LM, Mesh, Permutation_of_LM, boundary_types,
                  csr_Initialization, ebe_colour, gmres_mixed, ilu_levels, make_problem,
                  number_unknowns, pcg, pgmres, structured_mesh, symrcm)

__all__ = ["native", "CsrMatrix", "Device", "EbeMatrix", "Fill_LM", "Mesh", "Permutation_of_LM",
           "boundary_types", "csr_Initialization", "ebe_colour", "gmres_mixed", "ilu_levels",
           "make_problem", "number_unknowns", "pcg", "pgmres", "structured_mesh", "symrcm"]
