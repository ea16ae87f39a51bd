"""MI355X-native NebulaSEM hot path: dGSEM Euler residual and FV pressure PCG.

Python host mirror over the C-ABI in include/nebulasem_amd.h (ctypes); the
compute runs in libnebulasem_amd.so (hand-written gfx950 HIP kernels)."""
from ._lib import (BC_DIRICHLET, BC_NEUMANN, BC_ROBIN, BC_SYMMETRY, DGEuler, FVPoisson, Mesh,  # noqa: F401
                   NsemError, euler_params, grid_generate, lib, read_amr_tree, read_field, read_field_bc,
                   stream_copy, write_amr_tree, write_field, write_field_bc)
