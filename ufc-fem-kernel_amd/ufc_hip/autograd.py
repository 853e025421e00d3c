"""torch.autograd.Function over the C-ABI: a static solve u = K(E)^-1 f with a
hand-written adjoint.

forward : assemble K(E) on the GPU (ufc_rt_asm_assemble_k) and Jacobi-PCG
          (ufc_nm_pcg, FP64 or FP32-in-FP64 refinement).
backward: adjoint solve K^T lam = dL/du reusing the assembled operator
          (ufc_nm_adjoint_solve, the role of NM_Adjoint_Solve,
          NM_Solv_GMRESTranspose.f90:229-291 in the reference), then the element
          sensitivity kernel dL/dE_e = -lam_e^T K_e(E=1) u_e
          (ufc_ph_c3d8_sensitivity), which recomputes K_e instead of storing it.
Chain rules through E(rho) (e.g. SIMP E = E_min + rho^3 E_0) are left to torch.

Penalty Dirichlet rows and the loads do not depend on E, so they drop out of
dK/dE and df/dE.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _abi as A
from .static import Context


class StaticModel:
    """A mesh with loads and penalty BCs on one GPU context, solved for a
    per-element Young's modulus field E (a CUDA float64 tensor)."""

    def __init__(self, coords, conn, nu, bc_eq, bc_val=None, load=None, penalty=1.0e12,
                 precision=A.PREC_FP64, rtol=1.0e-10, inner_rtol=1.0e-5, adjoint_rtol=None,
                 device=0):
        self.ctx = Context(device)
        self.ctx.build_pattern(coords, conn)
        self.nu = float(nu)
        self.precision = precision
        self.rtol = rtol
        self.inner_rtol = inner_rtol
        self.adjoint_rtol = rtol if adjoint_rtol is None else adjoint_rtol
        self.device = torch.device("cuda", device)
        if load is not None:
            kind, elem, face, value = load
            if kind == "pressure":
                self.ctx.load_pressure(elem, face, value)
            else:
                self.ctx.load_traction(elem, face, value)
        self.ctx.apply_bc(bc_eq, bc_val, penalty)
        self.f = torch.as_tensor(self.ctx.get_rhs(), device=self.device)
        self.n_dof = self.ctx.n_dof
        self.n_elems = self.ctx.n_elems
        self._key = None
        self._young = None
        self.last_forward = None
        self.last_adjoint = None

    @classmethod
    def from_problem(cls, pr, **kw):
        load = None
        if pr.load_elem.size:
            load = (pr.load_kind, pr.load_elem, pr.load_face,
                    pr.load_value[0] if pr.load_kind == "pressure" else pr.load_value)
        return cls(pr.coords, pr.conn, pr.nu, pr.bc_eq, pr.bc_val, load, **kw)

    def assemble(self, young: torch.Tensor, reuse: bool = False):
        """Assemble K(young); with reuse, skip it when the operator on the
        device already is K(young).  The model keeps a reference to the field
        it assembled, so no other tensor can occupy its storage meanwhile."""
        key = (young.data_ptr(), young._version)
        if reuse and self._key == key and self._young is not None:
            return
        self.ctx.assemble_k(young.detach(), self.nu, precision=self.precision)
        self._young = young
        self._key = key

    def solve(self, young: torch.Tensor) -> torch.Tensor:
        return StaticSolve.apply(young, self)

    def close(self):
        self.ctx.close()


class StaticSolve(torch.autograd.Function):
    """u = K(E)^-1 f; gradient w.r.t. E by the adjoint method."""

    @staticmethod
    def forward(ctx, young: torch.Tensor, model: StaticModel):
        assert young.is_cuda and young.dtype == torch.float64 and young.numel() == model.n_elems
        young = young.detach().contiguous()
        model.assemble(young, reuse=False)
        u = torch.zeros(model.n_dof, dtype=torch.float64, device=young.device)
        _, info = model.ctx.pcg(precision=model.precision, rtol=model.rtol,
                                inner_rtol=model.inner_rtol, x=u)
        if not info.converged:
            raise A.UfcError(info.status, "forward PCG did not converge", "StaticSolve")
        model.last_forward = info
        ctx.model = model
        ctx.save_for_backward(young, u)
        return u

    @staticmethod
    def backward(ctx, grad_u: torch.Tensor):
        young, u = ctx.saved_tensors
        model: StaticModel = ctx.model
        model.assemble(young, reuse=True)  # the operator of the forward solve
        lam = torch.zeros_like(u)
        _, info = model.ctx.adjoint_solve(grad_u.detach().contiguous(), precision=model.precision,
                                          rtol=model.adjoint_rtol, inner_rtol=model.inner_rtol,
                                          x=lam)
        if not info.converged:
            raise A.UfcError(info.status, "adjoint PCG did not converge", "StaticSolve")
        model.last_adjoint = info
        grad = torch.empty_like(young)
        ones = torch.ones_like(young)
        model.ctx.sensitivity(u, lam, ones, model.nu, out=grad)
        return grad, None


def compliance(model: StaticModel, young: torch.Tensor, f_ext: torch.Tensor = None):
    """c = f^T u (f the external load: the penalty part of F is zero for
    homogeneous clamps)."""
    u = model.solve(young)
    f = model.f if f_ext is None else f_ext
    return torch.dot(f, u), u
