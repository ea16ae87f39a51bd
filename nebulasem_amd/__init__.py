"""MI355X-native NebulaSEM hot path: dGSEM Euler residual and FV pressure PCG.

Python host mirror over the C-ABI in include/nebulasem_amd.h (ctypes); the
compute runs in libnebulasem_amd.so (hand-written gfx950 HIP kernels)."""
from ._lib import (BC_DIRICHLET, BC_NEUMANN, BC_SYMMETRY, DGEuler, FVPoisson, Mesh, NsemError,  # noqa: F401
                   euler_params, grid_generate, lib, read_amr_tree, read_field, stream_copy, write_amr_tree,
                   write_field)
