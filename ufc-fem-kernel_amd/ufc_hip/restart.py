"""Restart files through ufc_rt_restart_save / ufc_rt_restart_load.

The file is the record sequence of RT_Out_RestartSave
(L5_RT/Output/RT_Out_Restart.f90:189-306) as Fortran unformatted sequential
records, optionally followed by the committed J2 state (13 doubles per Gauss
point: stress(6), eq. plastic strain, plastic strain(6)).  save()/load() with
ctx=None write and read u from host memory without a device; with a Context
they take u from host or device memory and can carry the J2 state.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
import os

import numpy as np

from . import _abi as A


@dataclass
class StepState:
    time: float = 0.0
    increment: int = 0
    lambda_: float = 0.0
    residual_norm: float = 0.0
    iterations: int = 0
    converged: bool = True


def save(path, u, step: StepState = StepState(), ctx=None, with_state=False):
    """Writes u (n_dof doubles) and the step data; with_state needs ctx."""
    u = np.ascontiguousarray(u, np.float64)
    d = A.RestartData(step.time, step.lambda_, step.residual_norm, int(step.increment),
                      int(step.iterations), int(bool(step.converged)), int(u.size),
                      u.ctypes.data_as(C.c_void_p))
    st = A.Status()
    h = ctx._h if ctx is not None else None
    A.check(A.lib().ufc_rt_restart_save(h, os.fsencode(path), C.byref(d), int(with_state),
                                        C.byref(st)), st)


def load(path, n_dof=None, ctx=None, with_state=False):
    """Reads a restart file: returns (u, StepState).  Without ctx, n_dof must
    be given and equal the file's; with ctx the context's mesh fixes it and
    with_state loads the J2 state into ctx."""
    if ctx is not None:
        n_dof = ctx.n_dof
    if not n_dof:
        raise ValueError("n_dof is needed without a context")
    u = np.empty(int(n_dof))
    d = A.RestartData(0.0, 0.0, 0.0, 0, 0, 0, int(n_dof), u.ctypes.data_as(C.c_void_p))
    st = A.Status()
    h = ctx._h if ctx is not None else None
    A.check(A.lib().ufc_rt_restart_load(h, os.fsencode(path), C.byref(d), int(with_state),
                                        C.byref(st)), st)
    return u, StepState(d.time, d.increment, d.lambda_, d.residual_norm, d.iterations,
                        bool(d.converged))
