"""ctypes mirror of include/ufc_hip.h (the C-ABI of libufc_hip.so).

The struct layouts below follow the header field by field; tests check that
every function the header declares is exported by the library.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("UFC_HIP_LIB") or os.path.join(PKG_DIR, "libufc_hip.so")

# status codes: IF_STATUS_* (L1_IF/Error/IF_Err_Brg.f90:95-109)
STATUS_OK = 0
STATUS_INVALID = 1
STATUS_ERROR = 2
STATUS_CONVERGED = 3
STATUS_NOT_CONVERGED = 4
STATUS_WARN = 6
STATUS_MEM_ERROR = 7
STATUS_FATAL = 10
STATUS_UNSUPPORTED = 11

PREC_FP64 = 0
PREC_FP32 = 1
PREC_MIXED = 2

KRYLOV_PCG = 0
KRYLOV_GMRES = 1

FP32_VIEW_ID32 = 0
FP32_VIEW_OFF16 = 1
FP32_VIEW_PATTERN = 2
FP32_VIEW_SYMDIAG = 3
# FP64 SpMV layouts (ufc_rt_asm_set_fp64_view)
FP64_VIEW_CSR = 0
FP64_VIEW_BLOCK = 1

MAT_ELASTIC = 0
MAT_J2 = 1

# C3D8 variants (PH_ELEM_C3D8_VARIANT_*, PH_Elem_C3D8.f90:111-116)
VARIANT_STANDARD = 1
VARIANT_REDUCED = 2
VARIANT_SELECTIVE = 3
VARIANT_INCOMPAT = 4
VARIANT_27PT = 6

# J2 hardening laws (PH_MAT_J2_HARD_*, PH_Mat_Plast_J2_Iso_Core.f90:33-36)
J2_HARD_LINEAR = 1
J2_HARD_SWIFT = 2
J2_HARD_VOCE = 3
J2_HARD_AF = 4

LOAD_TRACTION = 0
LOAD_PRESSURE = 1


class Status(C.Structure):
    _fields_ = [("status_code", C.c_int32), ("severity", C.c_int32), ("category", C.c_int32),
                ("line_number", C.c_int32), ("message", C.c_char * 512),
                ("source", C.c_char * 64)]


class MeshDesc(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("n_elems", C.c_int32), ("index_base", C.c_int32),
                ("reserved", C.c_int32), ("coords", C.c_void_p), ("conn", C.c_void_p)]


class AsmInfo(C.Structure):
    _fields_ = [("n_dof", C.c_int64), ("nnz", C.c_int64), ("n_nodes", C.c_int32),
                ("n_elems", C.c_int32), ("max_row_nnz", C.c_int32),
                ("nb_index_bytes", C.c_int32), ("fp32_view", C.c_int32),
                ("sym_tile", C.c_int32)]


class KeAlgo(C.Structure):
    _fields_ = [("precision", C.c_int32), ("variant", C.c_int32)]


class KeArgs(C.Structure):
    _fields_ = [("n_elems", C.c_int32), ("reserved", C.c_int32), ("coords", C.c_void_p),
                ("young", C.c_void_p), ("nu", C.c_double), ("ke", C.c_void_p)]


class SensArgs(C.Structure):
    _fields_ = [("u", C.c_void_p), ("lam", C.c_void_p), ("dedrho", C.c_void_p),
                ("nu", C.c_double), ("grad", C.c_void_p)]


class J2Props(C.Structure):
    _fields_ = [("young", C.c_double), ("nu", C.c_double), ("sigma_y0", C.c_double),
                ("hardening", C.c_double), ("hardening_type", C.c_int32),
                ("max_local_iter", C.c_int32), ("swift_k", C.c_double),
                ("swift_n", C.c_double), ("swift_eps0", C.c_double), ("voce_q", C.c_double),
                ("voce_delta", C.c_double), ("af_c", C.c_double), ("af_gamma", C.c_double)]


class AsmAlgo(C.Structure):
    _fields_ = [("precision", C.c_int32), ("material", C.c_int32), ("variant", C.c_int32),
                ("reserved", C.c_int32)]


class AsmArgs(C.Structure):
    _fields_ = [("young", C.c_void_p), ("nu", C.c_double)]


class BcArgs(C.Structure):
    _fields_ = [("n_bc", C.c_int32), ("index_base", C.c_int32), ("eq", C.c_void_p),
                ("value", C.c_void_p), ("penalty", C.c_double)]


class LoadArgs(C.Structure):
    _fields_ = [("n_faces", C.c_int32), ("kind", C.c_int32), ("index_base", C.c_int32),
                ("accumulate", C.c_int32), ("elem", C.c_void_p), ("face", C.c_void_p),
                ("magnitude", C.c_double), ("traction", C.c_double * 3)]


class SolverAlgo(C.Structure):
    _fields_ = [("precision", C.c_int32), ("max_iter", C.c_int32), ("rtol", C.c_double),
                ("inner_rtol", C.c_double), ("max_outer", C.c_int32),
                ("check_every", C.c_int32), ("replace_rtol", C.c_double),
                ("krylov", C.c_int32), ("restart", C.c_int32)]


class SolverState(C.Structure):
    _fields_ = [("x", C.c_void_p), ("b", C.c_void_p), ("niter", C.c_int32),
                ("n_outer", C.c_int32), ("rnorm", C.c_double), ("converged", C.c_int32),
                ("x0_zero", C.c_int32)]


class Timing(C.Structure):
    _fields_ = [("assemble_ms", C.c_double), ("solve_ms", C.c_double),
                ("spmv_ms_avg", C.c_double), ("spmv_calls", C.c_int32),
                ("n_replace", C.c_int32), ("j2_update_ms", C.c_double)]


class ApModel(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("n_elems", C.c_int32),
                ("node_label", C.c_void_p), ("elem_label", C.c_void_p),
                ("coords", C.c_void_p), ("conn", C.c_void_p), ("young", C.c_void_p),
                ("nu", C.c_double), ("n_bc", C.c_int32), ("n_cload", C.c_int32),
                ("bc_eq", C.c_void_p), ("bc_val", C.c_void_p),
                ("cload_eq", C.c_void_p), ("cload_val", C.c_void_p),
                ("n_face", C.c_int32), ("variant", C.c_int32),
                ("face_elem", C.c_void_p), ("face_id", C.c_void_p), ("face_p", C.c_void_p),
                ("impl", C.c_void_p)]


class RestartData(C.Structure):
    _fields_ = [("time", C.c_double), ("lambda_", C.c_double), ("residual_norm", C.c_double),
                ("increment", C.c_int32), ("iterations", C.c_int32), ("converged", C.c_int32),
                ("n_dof", C.c_int32), ("u", C.c_void_p)]


P = C.POINTER
_SIGS = {
    "ufc_ctx_create": [C.c_int32, P(C.c_void_p), P(Status)],
    "ufc_ctx_destroy": [C.c_void_p, P(Status)],
    "ufc_ctx_info": [C.c_void_p, P(AsmInfo), P(Status)],
    "ufc_ctx_timing": [C.c_void_p, P(Timing), P(Status)],
    "ufc_ctx_enable_timing": [C.c_void_p, C.c_int32, P(Status)],
    "ufc_ph_c3d8_eval_ke": [C.c_void_p, P(KeAlgo), P(KeArgs), P(Status)],
    "ufc_ph_c3d8_eval_ke_noalloc": [C.c_void_p, P(KeAlgo), P(KeArgs), P(Status)],
    "ufc_ph_c3d8_sensitivity": [C.c_void_p, P(SensArgs), P(Status)],
    "ufc_ph_j2_init": [C.c_void_p, P(J2Props), P(Status)],
    "ufc_ph_j2_update": [C.c_void_p, C.c_void_p, C.c_void_p, P(C.c_int64), P(Status)],
    "ufc_ph_j2_commit": [C.c_void_p, P(Status)],
    "ufc_ph_j2_get_state": [C.c_void_p, C.c_void_p, P(Status)],
    "ufc_ph_j2_set_state": [C.c_void_p, C.c_void_p, P(Status)],
    "ufc_rt_asm_build_pattern": [C.c_void_p, P(MeshDesc), P(AsmInfo), P(Status)],
    "ufc_rt_asm_get_pattern": [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, P(Status)],
    "ufc_rt_asm_assemble_k": [C.c_void_p, P(AsmAlgo), P(AsmArgs), P(Status)],
    "ufc_rt_asm_get_values": [C.c_void_p, C.c_int32, C.c_void_p, P(Status)],
    "ufc_rt_asm_apply_bc": [C.c_void_p, P(BcArgs), P(Status)],
    "ufc_rt_asm_load": [C.c_void_p, P(LoadArgs), P(Status)],
    "ufc_rt_set_rhs": [C.c_void_p, C.c_void_p, P(Status)],
    "ufc_rt_get_rhs": [C.c_void_p, C.c_void_p, P(Status)],
    "ufc_rt_residual": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, P(Status)],
    "ufc_nm_spmv": [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, P(Status)],
    "ufc_nm_pcg": [C.c_void_p, P(SolverAlgo), P(SolverState), P(Status)],
    "ufc_nm_spmv_t": [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, P(Status)],
    "ufc_rt_asm_set_values": [C.c_void_p, C.c_int32, C.c_void_p, P(Status)],
    "ufc_rt_asm_set_fp32_view": [C.c_void_p, C.c_int32, P(Status)],
    "ufc_rt_asm_set_fp64_view": [C.c_void_p, C.c_int32, P(Status)],
    "ufc_nm_adjoint_solve": [C.c_void_p, P(SolverAlgo), C.c_void_p, P(SolverState), P(Status)],
    "ufc_mesh_box": [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                     C.c_double, C.c_uint64, C.c_int32, C.c_void_p, C.c_void_p, P(Status)],
    "ufc_field": [C.c_int32, C.c_int64, C.c_uint64, C.c_double, C.c_double, C.c_void_p,
                  P(Status)],
    "ufc_mesh_box_numbering": [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                               P(Status)],
    "ufc_rt_restart_save": [C.c_void_p, C.c_char_p, P(RestartData), C.c_int32, P(Status)],
    "ufc_rt_restart_load": [C.c_void_p, C.c_char_p, P(RestartData), C.c_int32, P(Status)],
    "ufc_ap_read_inp": [C.c_char_p, P(ApModel), P(Status)],
    "ufc_ap_free": [P(ApModel), P(Status)],
}

_lib = None


class UfcError(RuntimeError):
    """A non-OK ufc_status; .code is the IF_STATUS_* value."""

    def __init__(self, code: int, message: str, source: str):
        super().__init__(f"[{source}] status {code}: {message}")
        self.code = code
        self.message = message
        self.source = source


def lib():
    """The loaded libufc_hip.so.  Raises if it has not been built: there is no
    fallback path."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(or python ufc-fem-kernel_amd/ufc_hip/build.py)")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def check(code: int, st: Status, ok=(STATUS_OK,)):
    if code not in ok:
        raise UfcError(code, st.message.decode(errors="replace"),
                       st.source.decode(errors="replace"))
    return code
