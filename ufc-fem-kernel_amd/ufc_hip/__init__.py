"""ufc_hip -- MI355X (gfx950) static-stiffness path for UFC over libufc_hip.so.

C3D8 element stiffness, atomics-free fixed-pattern CSR assembly, CSR SpMV and
Jacobi-PCG (FP64, FP32, FP32-in-FP64 refinement), adjoint sensitivities and
the J2 material update, behind the C-ABI of include/ufc_hip.h.
"""
from ._abi import (PREC_FP64, PREC_FP32, PREC_MIXED, MAT_ELASTIC, MAT_J2, UfcError,  # noqa
                   STATUS_OK, STATUS_INVALID, STATUS_ERROR, lib, LIB_PATH)
from .static import Context, SolveInfo, mesh_box, field  # noqa

__all__ = ["Context", "SolveInfo", "mesh_box", "field", "UfcError", "PREC_FP64", "PREC_FP32",
           "PREC_MIXED", "MAT_ELASTIC", "MAT_J2", "lib", "LIB_PATH"]
