"""Host-side mirror of UFC's static-stiffness interfaces over libufc_hip.so.

`Context` exposes the C-ABI bundles with the reference's names and argument
meaning (reference paths relative to its root):

  build_pattern   RT_Asm_GlobalStiffness full-build path (RT_Asm_Solv.f90:1902-1965)
                  + RT_CSR_FromTripletMerged (RT_Solv_Sparse.f90:336-447)
  assemble_k      RT_Asm_GlobalStiffness reuse path (RT_Asm_Solv.f90:1840-1900)
  apply_bc        RT_Asm_ApplyBC _Idx path (RT_Asm_Solv.f90:318-344)
  load_traction / load_pressure
                  RT_Asm_GlobalLoad surface branch (RT_Asm_Solv.f90:1614-1647)
  eval_ke         PH_Elem_Eval_Ke -> PH_Elem_C3D8_StiffMatrix (PH_Elem_C3D8.f90:759-780)
  spmv            NM_CSR_MatVec_Optimized (NM_Mtx_Sparse.f90:341-365)
  pcg             NM_Solver_PCG (NM_Solv_Core.f90:363-423)
  adjoint_solve   NM_Adjoint_Solve (NM_Solv_GMRESTranspose.f90:229-291)
  sensitivity     new: -dE/drho lam_e^T K_e(E=1) u_e
  j2_*            PH_J2_ComputeStress (PH_Mat_Plast_J2_Iso_Core.f90:133-207) on every Gauss point
  residual        RT_Asm_ComputeResidual (RT_Asm_Solv.f90:1311-1426)

Array arguments may be numpy arrays (host) or torch tensors on the context's
GPU; every call is blocking.  Errors raise `UfcError` carrying the IF_STATUS
code, as the reference's ErrorStatusType does; nothing falls back to a CPU
path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi as A

try:  # torch is optional for host-array use
    import torch
except Exception:  # pragma: no cover
    torch = None


def _is_tensor(a) -> bool:
    return torch is not None and isinstance(a, torch.Tensor)


class _Arg:
    """Keeps converted arrays alive for the duration of a call."""

    def __init__(self):
        self.keep = []

    def ptr(self, a, dtype):
        if a is None:
            return None
        if _is_tensor(a):
            tdt = {np.float64: torch.float64, np.float32: torch.float32,
                   np.int32: torch.int32}[dtype]
            t = a.detach()
            if t.dtype != tdt or not t.is_contiguous():
                t = t.to(tdt).contiguous()
            if t.is_cuda:
                torch.cuda.current_stream(t.device).synchronize()
            self.keep.append(t)
            return C.c_void_p(t.data_ptr())
        arr = np.ascontiguousarray(a, dtype=dtype)
        self.keep.append(arr)
        return arr.ctypes.data_as(C.c_void_p)

    def out(self, a, dtype):
        """Output array: must already have the right dtype and be contiguous."""
        if _is_tensor(a):
            assert a.is_contiguous(), "output tensor must be contiguous"
            if a.is_cuda:
                torch.cuda.current_stream(a.device).synchronize()
            return C.c_void_p(a.data_ptr())
        assert isinstance(a, np.ndarray) and a.dtype == dtype and a.flags.c_contiguous
        return a.ctypes.data_as(C.c_void_p)


@dataclass
class SolveInfo:
    status: int
    niter: int
    n_outer: int
    rnorm: float
    converged: bool
    solve_ms: float
    spmv_ms_avg: float
    spmv_calls: int


def mesh_box(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, jitter=0.0, seed=0, index_base=0):
    """Structured C3D8 box from the library's generator: coords (n,3), conn (e,8)."""
    L = A.lib()
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    ne = nx * ny * nz
    coords = np.empty((nn, 3), np.float64)
    conn = np.empty((ne, 8), np.int32)
    st = A.Status()
    A.check(L.ufc_mesh_box(nx, ny, nz, lx, ly, lz, jitter, seed, index_base,
                           coords.ctypes.data_as(C.c_void_p), conn.ctypes.data_as(C.c_void_p),
                           C.byref(st)), st)
    return coords, conn


def box_numbering(conn, n_nodes, index_base=0):
    """ufc_mesh_box_numbering: perm (n_nodes,) with perm[n] the lexicographic
    number of node n when the hex mesh is a structured box under some node
    numbering, else None."""
    L = A.lib()
    conn = np.ascontiguousarray(conn, np.int32)
    perm = np.empty(n_nodes, np.int32)
    st = A.Status()
    code = L.ufc_mesh_box_numbering(n_nodes, conn.shape[0], index_base,
                                    conn.ctypes.data_as(C.c_void_p),
                                    perm.ctypes.data_as(C.c_void_p), C.byref(st))
    if code == A.STATUS_UNSUPPORTED:
        return None
    A.check(code, st)
    return perm


def field(kind: str, n: int, seed: int, a: float, b: float) -> np.ndarray:
    """kind 'lognormal': exp(a + b z); kind 'uniform': U[a, b]; keyed by (seed, element)."""
    L = A.lib()
    out = np.empty(n, np.float64)
    st = A.Status()
    k = {"lognormal": 0, "uniform": 1}[kind]
    A.check(L.ufc_field(k, n, seed, a, b, out.ctypes.data_as(C.c_void_p), C.byref(st)), st)
    return out


class Context:
    """One GPU context (ufc_ctx) holding a mesh, its CSR pattern, assembled
    values, loads, BCs, solver work arrays and optional J2 state."""

    def __init__(self, device: int = 0):
        self._L = A.lib()
        self._h = C.c_void_p()
        st = A.Status()
        A.check(self._L.ufc_ctx_create(device, C.byref(self._h), C.byref(st)), st)
        self.device = device
        self.n_dof = 0
        self.nnz = 0
        self.n_nodes = 0
        self.n_elems = 0

    # -------------------------------------------------------------- lifetime
    def close(self):
        if self._h:
            st = A.Status()
            self._L.ufc_ctx_destroy(self._h, C.byref(st))
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, fname, *args, ok=(A.STATUS_OK,)):
        st = A.Status()
        code = getattr(self._L, fname)(self._h, *args, C.byref(st))
        return A.check(code, st, ok)

    # -------------------------------------------------------------- L5 pattern / assembly
    def build_pattern(self, coords, conn, index_base=0):
        coords = np.ascontiguousarray(coords, np.float64)
        conn = np.ascontiguousarray(conn, np.int32)
        m = A.MeshDesc(coords.shape[0], conn.shape[0], index_base, 0,
                       coords.ctypes.data_as(C.c_void_p), conn.ctypes.data_as(C.c_void_p))
        info = A.AsmInfo()
        self._call("ufc_rt_asm_build_pattern", C.byref(m), C.byref(info))
        self.n_dof, self.nnz = info.n_dof, info.nnz
        self.n_nodes, self.n_elems = info.n_nodes, info.n_elems
        self.nb_index_bytes = info.nb_index_bytes
        self.fp32_view = info.fp32_view
        self.sym_tile = info.sym_tile
        return info

    def set_fp32_view(self, view):
        """Select the FP32 SpMV view (A.FP32_VIEW_*, or -1 for the fastest the
        mesh allows); raises on UFC_STATUS_UNSUPPORTED when the mesh does not
        allow it.  Updates fp32_view and nb_index_bytes."""
        self._call("ufc_rt_asm_set_fp32_view", int(view))
        i = self.info()
        self.fp32_view, self.nb_index_bytes = i.fp32_view, i.nb_index_bytes
        return self.fp32_view

    def set_fp64_view(self, view):
        """Layout of the FP64 SpMV: A.FP64_VIEW_CSR (a column index per value,
        the reference's) or A.FP64_VIEW_BLOCK (one neighbour id per 3x3
        block); both give the same bits."""
        self._call("ufc_rt_asm_set_fp64_view", int(view))

    def get_pattern(self, index_base=0):
        rp = np.empty(self.n_dof + 1, np.int32)
        ci = np.empty(self.nnz, np.int32)
        self._call("ufc_rt_asm_get_pattern", index_base, rp.ctypes.data_as(C.c_void_p),
                   ci.ctypes.data_as(C.c_void_p))
        return rp, ci

    def assemble_k(self, young, nu, precision=A.PREC_FP64, material=A.MAT_ELASTIC,
                   variant=A.VARIANT_STANDARD):
        arg = _Arg()
        algo = A.AsmAlgo(precision, material, variant, 0)
        args = A.AsmArgs(arg.ptr(young, np.float64) if young is not None else None, nu)
        self._call("ufc_rt_asm_assemble_k", C.byref(algo), C.byref(args))

    def get_values(self, precision=A.PREC_FP64):
        dt = np.float64 if precision == A.PREC_FP64 else np.float32
        v = np.empty(self.nnz, dt)
        self._call("ufc_rt_asm_get_values", precision, v.ctypes.data_as(C.c_void_p))
        return v

    def apply_bc(self, eq, value=None, penalty=1.0e12, index_base=0):
        arg = _Arg()
        eq = np.ascontiguousarray(eq, np.int32)
        a = A.BcArgs(eq.size, index_base, arg.ptr(eq, np.int32),
                     arg.ptr(value, np.float64) if value is not None else None, penalty)
        self._call("ufc_rt_asm_apply_bc", C.byref(a))

    def _load(self, kind, elem, face, magnitude=0.0, traction=(0.0, 0.0, 0.0), index_base=0,
              accumulate=False):
        arg = _Arg()
        elem = np.ascontiguousarray(elem, np.int32)
        face = np.ascontiguousarray(face, np.int32)
        a = A.LoadArgs(elem.size, kind, index_base, int(accumulate), arg.ptr(elem, np.int32),
                       arg.ptr(face, np.int32), magnitude, (C.c_double * 3)(*traction))
        self._call("ufc_rt_asm_load", C.byref(a))

    def load_traction(self, elem, face, traction, index_base=0, accumulate=False):
        """traction x area/4 on each face node (MD_LBC_Apply.f90:258-293)."""
        self._load(A.LOAD_TRACTION, elem, face, traction=tuple(traction), index_base=index_base,
                   accumulate=accumulate)

    def load_pressure(self, elem, face, p, index_base=0, accumulate=False):
        """consistent face pressure along the outward normal (PH_Elem_C3D8.f90:2360-2541)."""
        self._load(A.LOAD_PRESSURE, elem, face, magnitude=p, index_base=index_base,
                   accumulate=accumulate)

    def set_rhs(self, f):
        arg = _Arg()
        self._call("ufc_rt_set_rhs", arg.ptr(f, np.float64))

    def get_rhs(self, out=None):
        out = np.empty(self.n_dof) if out is None else out
        arg = _Arg()
        self._call("ufc_rt_get_rhs", arg.out(out, np.float64))
        return out

    def residual(self, u, out=None):
        arg = _Arg()
        out = np.empty(self.n_dof) if out is None else out
        norms = np.empty(2)
        self._call("ufc_rt_residual", arg.ptr(u, np.float64), arg.out(out, np.float64),
                   norms.ctypes.data_as(C.c_void_p))
        return out, norms

    # -------------------------------------------------------------- L4 element kernels
    def eval_ke(self, xe, young, nu, precision=A.PREC_FP64, variant=A.VARIANT_STANDARD):
        """Batched Ke: xe (n,8,3), young (n,) -> (n,24,24) of the C3D8 variant."""
        xe = np.ascontiguousarray(xe, np.float64)
        young = np.ascontiguousarray(young, np.float64)
        n = xe.shape[0]
        dt = np.float64 if precision == A.PREC_FP64 else np.float32
        ke = np.empty((n, 24, 24), dt)
        algo = A.KeAlgo(precision, variant)
        args = A.KeArgs(n, 0, xe.ctypes.data_as(C.c_void_p), young.ctypes.data_as(C.c_void_p),
                        nu, ke.ctypes.data_as(C.c_void_p))
        self._call("ufc_ph_c3d8_eval_ke", C.byref(algo), C.byref(args))
        return np.transpose(ke, (0, 2, 1))  # column-major per element

    def eval_ke_noalloc(self, xe, young, nu, ke, precision=A.PREC_FP64,
                        variant=A.VARIANT_STANDARD):
        """ufc_ph_c3d8_eval_ke_noalloc on device tensors: xe (n,8,3) float64,
        young (n,) float64 and ke (n,24,24) float64/float32, all contiguous on
        the context's GPU; Ke is written column-major per element, as
        eval_ke's buffer.  Allocates nothing on the device."""
        for t in (xe, young, ke):
            assert _is_tensor(t) and t.is_cuda and t.is_contiguous(), \
                "eval_ke_noalloc takes contiguous device tensors"
        torch.cuda.current_stream(ke.device).synchronize()
        algo = A.KeAlgo(precision, variant)
        args = A.KeArgs(int(xe.shape[0]), 0, C.c_void_p(xe.data_ptr()),
                        C.c_void_p(young.data_ptr()), nu, C.c_void_p(ke.data_ptr()))
        self._call("ufc_ph_c3d8_eval_ke_noalloc", C.byref(algo), C.byref(args))
        return ke

    def sensitivity(self, u, lam, dedrho, nu, out=None):
        arg = _Arg()
        out = np.empty(self.n_elems) if out is None else out
        a = A.SensArgs(arg.ptr(u, np.float64), arg.ptr(lam, np.float64),
                       arg.ptr(dedrho, np.float64), nu, arg.out(out, np.float64))
        self._call("ufc_ph_c3d8_sensitivity", C.byref(a))
        return out

    def j2_init(self, young, nu, sigma_y0, hardening=0.0, hardening_type=A.J2_HARD_LINEAR,
                max_local_iter=0, swift_k=0.0, swift_n=0.0, swift_eps0=0.0, voce_q=0.0,
                voce_delta=0.0, af_c=0.0, af_gamma=0.0):
        """ufc_ph_j2_init: zero J2 state with the hardening law of
        PH_J2_ComputeHardening (linear H, Swift K (eps0 + e)^n or Voce
        sigma_y0 + Q (1 - exp(-delta e))); max_local_iter bounds the local
        Newton (0: 25)."""
        p = A.J2Props(young, nu, sigma_y0, hardening, int(hardening_type), int(max_local_iter),
                      swift_k, swift_n, swift_eps0, voce_q, voce_delta, af_c, af_gamma)
        self._call("ufc_ph_j2_init", C.byref(p))

    def j2_update(self, u, u_n):
        arg = _Arg()
        ny = C.c_int64()
        self._call("ufc_ph_j2_update", arg.ptr(u, np.float64), arg.ptr(u_n, np.float64),
                   C.byref(ny))
        return ny.value

    def j2_commit(self):
        self._call("ufc_ph_j2_commit")

    def j2_get_state(self):
        out = np.empty((self.n_elems, 8, 13))
        self._call("ufc_ph_j2_get_state", out.ctypes.data_as(C.c_void_p))
        return out

    def j2_set_state(self, state):
        arg = _Arg()
        self._call("ufc_ph_j2_set_state", arg.ptr(state, np.float64))

    # -------------------------------------------------------------- L2 solvers
    def spmv(self, x, precision=A.PREC_FP64, out=None):
        arg = _Arg()
        out = np.empty(self.n_dof) if out is None else out
        self._call("ufc_nm_spmv", precision, arg.ptr(x, np.float64), arg.out(out, np.float64))
        return out

    def enable_timing(self, per_kernel=True):
        """HIP event pairs around every per_kernel-th SpMV launch (True: every
        launch, 0/False: none); timing().spmv_ms_avg is their mean."""
        self._call("ufc_ctx_enable_timing", int(per_kernel))

    def timing(self):
        t = A.Timing()
        self._call("ufc_ctx_timing", C.byref(t))
        return t

    def _solve(self, fname, extra, precision, rtol, max_iter, inner_rtol, max_outer, x, b,
               x0, check_every, replace_rtol, krylov=A.KRYLOV_PCG, restart=30):
        arg = _Arg()
        out = np.empty(self.n_dof) if x is None else x
        xp = arg.out(out, np.float64)
        x0_zero = 1
        if x0 is not None:
            # the solution buffer doubles as the initial guess
            if _is_tensor(out):
                out.copy_(x0 if _is_tensor(x0) else torch.as_tensor(x0, device=out.device))
            else:
                out[...] = np.asarray(x0)
            x0_zero = 0
        algo = A.SolverAlgo(precision, max_iter, rtol, inner_rtol, max_outer, check_every,
                            replace_rtol, krylov, restart)
        s = A.SolverState(xp, arg.ptr(b, np.float64) if b is not None else None, 0, 0, 0.0, 0,
                          x0_zero)
        args = [C.byref(algo)] + ([arg.ptr(extra, np.float64)] if extra is not None else []) + \
            [C.byref(s)]
        code = self._call(fname, *args,
                          ok=(A.STATUS_OK, A.STATUS_INVALID, A.STATUS_ERROR))
        t = self.timing()
        return out, SolveInfo(code, s.niter, s.n_outer, s.rnorm, bool(s.converged),
                              t.solve_ms, t.spmv_ms_avg, t.spmv_calls)

    def pcg(self, precision=A.PREC_FP64, rtol=1e-10, max_iter=100000, inner_rtol=1e-5,
            max_outer=20, x=None, b=None, x0=None, check_every=0, replace_rtol=0.0):
        """Jacobi-PCG on the assembled K; returns (x, SolveInfo).  status is
        OK, INVALID (max_iter reached, NM_Solv_Core.f90:419-422) or ERROR
        (breakdown).  MIXED: the inner FP32 PCG replaces its residual by the
        FP64 one each time it has dropped by replace_rtol, keeping the search
        direction (<= 0: 0.1); replace_rtol >= 1 restarts the FP32 PCG at each
        refinement step instead."""
        return self._solve("ufc_nm_pcg", None, precision, rtol, max_iter, inner_rtol,
                           max_outer, x, b, x0, check_every, replace_rtol)

    def adjoint_solve(self, dJdu, precision=A.PREC_FP64, rtol=1e-10, max_iter=100000,
                      inner_rtol=1e-5, max_outer=20, x=None, check_every=0, replace_rtol=0.0,
                      krylov=A.KRYLOV_PCG, restart=30):
        """K^T lambda = dJdu: Jacobi-PCG (K symmetric) or, krylov=KRYLOV_GMRES,
        GMRES(restart) on K^T over the FP64 values (info.n_outer = restarts)."""
        return self._solve("ufc_nm_adjoint_solve", dJdu, precision, rtol, max_iter, inner_rtol,
                           max_outer, x, None, None, check_every, replace_rtol, krylov, restart)

    def spmv_t(self, x, out=None):
        """y = K^T x with the FP64 values."""
        arg = _Arg()
        out = np.empty(self.n_dof) if out is None else out
        self._call("ufc_nm_spmv_t", A.PREC_FP64, arg.ptr(x, np.float64),
                   arg.out(out, np.float64))
        return out

    def set_values(self, values):
        """Replace the FP64 values (CSR order, penalties included, as
        get_values(PREC_FP64) returns them)."""
        arg = _Arg()
        self._call("ufc_rt_asm_set_values", A.PREC_FP64, arg.ptr(values, np.float64))

    def info(self):
        i = A.AsmInfo()
        self._call("ufc_ctx_info", C.byref(i))
        return i
