"""sgipc -- host-side Python mirror of the libsgipc.so C ABI (MI355X / gfx950
implementation of the Stiff-GIPC Newton-step hot path)."""
from . import scene  # noqa: F401
from ._lib import (LinearSystem, Simulation, SgipcError, check, device_info, lib,  # noqa: F401
                   read_mm_vector, read_mm_upper_bcoo, spatial_sort_keys,
                   SGIPC_PRECOND_DIAG, SGIPC_PRECOND_MAS, SGIPC_MAS_FP32, SGIPC_MAS_FP64,
                   SGIPC_SOLVE_FP64, SGIPC_SOLVE_MIXED, SGIPC_DOF_FEM_FIRST,
                   SGIPC_DOF_ABD_FIRST)
from .sort import SortResult, spatial_sort  # noqa: F401,E402
