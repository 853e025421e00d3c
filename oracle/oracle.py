"""ctypes/numpy front end of the FP64 CPU oracle (oracle/libufc_oracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline -- never as the thing measured or shipped.

Every wrapper names the C function it calls; the C file cites the reference
routines (paths relative to the reference root).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libufc_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.orc_rng_u64.restype = C.c_uint64
        L.orc_rng_u64.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_rng_u01.restype = C.c_double
        L.orc_rng_u01.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_field_lognormal.argtypes = [C.c_int64, C.c_uint64, C.c_double, C.c_double, vp]
        L.orc_field_uniform.argtypes = [C.c_int64, C.c_uint64, C.c_double, C.c_double, vp]
        L.orc_mesh_box.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_double, vp, vp]
        L.orc_mesh_jitter.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_uint64, vp]
        L.orc_c3d8_shape.argtypes = [C.c_double, C.c_double, C.c_double, vp, vp]
        L.orc_c3d8_dndx.restype = C.c_int
        L.orc_c3d8_dndx.argtypes = [vp, C.c_double, C.c_double, C.c_double, vp, vp]
        L.orc_c3d8_ke.restype = C.c_int
        L.orc_c3d8_ke.argtypes = [vp, C.c_double, C.c_double, vp]
        L.orc_c3d8_ke_variant.restype = C.c_int
        L.orc_c3d8_ke_variant.argtypes = [vp, C.c_double, C.c_double, C.c_int, vp]
        L.orc_c3d8_ke_dgp.restype = C.c_int
        L.orc_c3d8_ke_dgp.argtypes = [vp, vp, vp]
        L.orc_c3d8_strain.argtypes = [vp, vp, vp]
        L.orc_c3d8_fint.argtypes = [vp, vp, vp]
        L.orc_c3d8_face_pressure.argtypes = [vp, C.c_int, C.c_double, vp]
        L.orc_face_area.restype = C.c_double
        L.orc_face_area.argtypes = [vp, C.c_int]
        L.orc_csr_pattern.restype = C.c_int64
        L.orc_csr_pattern.argtypes = [C.c_int32, C.c_int32, vp, vp, vp]
        L.orc_csr_from_triplets.restype = C.c_int64
        L.orc_csr_from_triplets.argtypes = [C.c_int32, C.c_int64, vp, vp, vp, vp, vp, vp]
        L.orc_assemble.restype = C.c_int
        L.orc_assemble.argtypes = [C.c_int32, vp, vp, vp, C.c_double, C.c_int32, vp, vp, vp,
                                   C.c_int, vp]
        L.orc_assemble_dgp.restype = C.c_int
        L.orc_assemble_dgp.argtypes = [C.c_int32, vp, vp, vp, C.c_int32, vp, vp, vp, C.c_int]
        L.orc_apply_penalty.restype = C.c_int
        L.orc_apply_penalty.argtypes = [C.c_int32, vp, vp, C.c_double, vp, vp, vp, vp]
        L.orc_face_loads_traction.argtypes = [C.c_int32, vp, vp, vp, vp, vp, vp]
        L.orc_face_loads_pressure.argtypes = [C.c_int32, vp, vp, vp, vp, C.c_double, vp]
        L.orc_spmv.argtypes = [C.c_int32, vp, vp, vp, vp, vp, C.c_int]
        L.orc_spmv_t.restype = None
        L.orc_spmv_t.argtypes = [C.c_int32, vp, vp, vp, vp, vp]
        L.orc_pcg.restype = C.c_int
        L.orc_pcg.argtypes = [C.c_int32, vp, vp, vp, vp, vp, C.c_int32, C.c_double, vp, vp,
                              C.c_int]
        L.orc_pcg_mark.restype = C.c_int
        L.orc_pcg_mark.argtypes = [C.c_int32, vp, vp, vp, vp, vp, C.c_int32, C.c_double,
                                   C.c_double, vp, vp, vp, vp, C.c_int, C.c_int32]
        L.orc_cg_transpose.restype = C.c_int
        L.orc_cg_transpose.argtypes = [C.c_int32, vp, vp, vp, vp, vp, C.c_double, C.c_int32,
                                       vp, vp]
        L.orc_sensitivity.argtypes = [C.c_int32, vp, vp, C.c_double, vp, vp, vp, vp, C.c_int]
        L.orc_j2_internal_force.restype = C.c_int64
        L.orc_j2_internal_force.argtypes = [C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, vp,
                                            vp, vp, C.c_int]
        L.orc_j2_update.restype = C.c_int
        L.orc_j2_update.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.orc_j2_update_ex.restype = C.c_int
        L.orc_j2_update_ex.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_j2_newton_step.restype = C.c_int
        L.orc_j2_newton_step.argtypes = [C.c_int32, C.c_int32, vp, vp, vp, C.c_int32, vp, vp,
                                         C.c_double, vp, vp, vp, C.c_int32, C.c_double,
                                         C.c_double, vp, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


# ----------------------------------------------------------------- fields, mesh
def field_lognormal(n: int, seed: int, mu: float, sigma: float) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().orc_field_lognormal(n, seed, mu, sigma, _p(out))
    return out


def field_uniform(n: int, seed: int, a: float, b: float) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().orc_field_uniform(n, seed, a, b, _p(out))
    return out


def mesh_box(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, jitter=0.0, seed=0):
    """Structured C3D8 box: coords (n_nodes,3) and 0-based conn (n_elems,8)."""
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    ne = nx * ny * nz
    coords = np.empty((nn, 3), np.float64)
    conn = np.empty((ne, 8), np.int32)
    lib().orc_mesh_box(nx, ny, nz, lx, ly, lz, _p(coords), _p(conn))
    if jitter:
        lib().orc_mesh_jitter(nx, ny, nz, lx, ly, lz, jitter, seed, _p(coords))
    return coords, conn


# ----------------------------------------------------------------- element
def shape(xi, eta, zeta):
    N = np.empty(8)
    dN = np.empty((8, 3))
    lib().orc_c3d8_shape(xi, eta, zeta, _p(N), _p(dN))
    return N, dN


def dndx(xe, xi, eta, zeta):
    xe = _f64(xe)
    g = np.empty((8, 3))
    d = C.c_double()
    ok = lib().orc_c3d8_dndx(_p(xe), xi, eta, zeta, _p(g), C.byref(d))
    return g, d.value, ok


def ke(xe, E, nu) -> np.ndarray:
    """24x24 element stiffness (symmetric); xe is (8,3) node coordinates."""
    xe = _f64(xe)
    K = np.empty((24, 24))
    lib().orc_c3d8_ke(_p(xe), E, nu, _p(K))
    return K.T.copy()  # stored column-major


VARIANT_STANDARD, VARIANT_REDUCED, VARIANT_SELECTIVE, VARIANT_INCOMPAT, VARIANT_27PT = 1, 2, 3, 4, 6


def ke_variant(xe, E, nu, variant) -> np.ndarray:
    """24x24 stiffness of a C3D8 variant (PH_Elem_C3D8_StiffMatrixByVariant,
    PH_Elem_C3D8.f90:400-417): 1 standard, 2 reduced + hourglass, 3 selective
    (B-bar), 4 incompatible modes, 6 27-point."""
    xe = _f64(xe)
    K = np.empty((24, 24))
    rc = lib().orc_c3d8_ke_variant(_p(xe), E, nu, int(variant), _p(K))
    if rc < 0:
        raise ValueError(f"unknown variant {variant}")
    if rc == 0:
        raise ArithmeticError("incompatible-mode block not positive definite")
    return K.T.copy()


def ke_dgp(xe, Dgp) -> np.ndarray:
    """Ke with one 6x6 D per Gauss point; Dgp is (8,6,6) with Dgp[g][i][j]=D_ij."""
    xe = _f64(xe)
    Dc = _f64(np.transpose(np.asarray(Dgp, np.float64), (0, 2, 1)))  # column-major per GP
    K = np.empty((24, 24))
    lib().orc_c3d8_ke_dgp(_p(xe), _p(Dc), _p(K))
    return K.T.copy()


def strain(xe, ue) -> np.ndarray:
    out = np.empty((8, 6))
    lib().orc_c3d8_strain(_p(_f64(xe)), _p(_f64(ue)), _p(out))
    return out


def fint(xe, sig_gp) -> np.ndarray:
    out = np.empty(24)
    lib().orc_c3d8_fint(_p(_f64(xe)), _p(_f64(sig_gp)), _p(out))
    return out


def face_pressure(xe, face: int, p: float) -> np.ndarray:
    out = np.empty(24)
    lib().orc_c3d8_face_pressure(_p(_f64(xe)), face, p, _p(out))
    return out


def face_area(xe, face: int) -> float:
    return lib().orc_face_area(_p(_f64(xe)), face)


# ----------------------------------------------------------------- pattern
def csr_pattern(n_nodes: int, conn: np.ndarray):
    conn = _i32(conn)
    rp = np.empty(3 * n_nodes + 1, np.int32)
    nnz = lib().orc_csr_pattern(n_nodes, conn.shape[0], _p(conn), _p(rp), None)
    ci = np.empty(nnz, np.int32)
    lib().orc_csr_pattern(n_nodes, conn.shape[0], _p(conn), _p(rp), _p(ci))
    return rp, ci


def assemble_variant(conn, coords, E, nu, variant, rp, ci):
    """CSR values on (rp, ci) from the per-element ke_variant matrices, merged
    as RT_CSR_FromTripletMerged does (element order within each entry)."""
    conn = np.asarray(conn)
    n_rows = rp.size - 1
    ti, tj, tv = [], [], []
    for e in range(conn.shape[0]):
        K = ke_variant(coords[conn[e]], float(E[e]), nu, variant)
        d = (3 * conn[e][:, None] + np.arange(3)).ravel()
        ti.append(np.repeat(d, 24))
        tj.append(np.tile(d, 24))
        tv.append(K.ravel())
    rp2, ci2, val = csr_from_triplets(n_rows, np.concatenate(ti), np.concatenate(tj),
                                      np.concatenate(tv))
    if not (np.array_equal(rp2, rp) and np.array_equal(ci2, ci)):
        raise ValueError("pattern differs from the element triplets")
    return val


def csr_from_triplets(n_rows, ti, tj, tv=None):
    ti, tj = _i32(ti), _i32(tj)
    rp = np.empty(n_rows + 1, np.int32)
    nnz = lib().orc_csr_from_triplets(n_rows, ti.size, _p(ti), _p(tj), None, _p(rp), None, None)
    ci = np.empty(nnz, np.int32)
    val = np.empty(nnz)
    tvp = _p(_f64(tv)) if tv is not None else None
    tvk = _f64(tv) if tv is not None else None  # keep alive
    lib().orc_csr_from_triplets(n_rows, ti.size, _p(ti), _p(tj),
                                _p(tvk) if tvk is not None else tvp, _p(rp), _p(ci), _p(val))
    return rp, ci, val


def assemble(conn, coords, E, nu, rp, ci, nthreads=1, allow_degenerate=False):
    """CSR values of the elastic K on (rp, ci).  Elements with a Gauss point
    of |det J| <= 1e-12 are an error unless allow_degenerate, in which case
    the point is skipped as the reference skips it."""
    conn, coords, E = _i32(conn), _f64(coords), _f64(E)
    val = np.empty(ci.size)
    degen = C.c_int32(0)
    bad = lib().orc_assemble(conn.shape[0], _p(conn), _p(coords), _p(E), nu, rp.size - 1,
                             _p(rp), _p(ci), _p(val), nthreads, C.byref(degen))
    if bad:
        raise RuntimeError(f"orc_assemble: {bad} element entries without a pattern slot")
    if degen.value and not allow_degenerate:
        raise RuntimeError(f"orc_assemble: {degen.value} elements with degenerate Gauss points")
    return val


def assemble_dgp(conn, coords, Dgp_colmajor, rp, ci, nthreads=1):
    conn, coords, D = _i32(conn), _f64(coords), _f64(Dgp_colmajor)
    val = np.empty(ci.size)
    lib().orc_assemble_dgp(conn.shape[0], _p(conn), _p(coords), _p(D), rp.size - 1, _p(rp),
                           _p(ci), _p(val), nthreads)
    return val


def apply_penalty(eq, value, penalty, rp, ci, val, F):
    eq = _i32(eq)
    value = _f64(value)
    return lib().orc_apply_penalty(eq.size, _p(eq), _p(value), penalty, _p(rp), _p(ci),
                                   _p(val) if val is not None else None,
                                   _p(F) if F is not None else None)


def face_loads_traction(elem, face, conn, coords, traction, F):
    elem, face, conn, coords, t = _i32(elem), _i32(face), _i32(conn), _f64(coords), _f64(traction)
    lib().orc_face_loads_traction(elem.size, _p(elem), _p(face), _p(conn), _p(coords), _p(t),
                                  _p(F))


def face_loads_pressure(elem, face, conn, coords, p, F):
    elem, face, conn, coords = _i32(elem), _i32(face), _i32(conn), _f64(coords)
    lib().orc_face_loads_pressure(elem.size, _p(elem), _p(face), _p(conn), _p(coords), p, _p(F))


# ----------------------------------------------------------------- solvers
def spmv(rp, ci, val, x, nthreads=1):
    y = np.empty(rp.size - 1)
    lib().orc_spmv(rp.size - 1, _p(rp), _p(ci), _p(val), _p(_f64(x)), _p(y), nthreads)
    return y


def spmv_t(rp, ci, val, x):
    """y = A^T x (NM_CSR_Transpose + SpMV, NM_Mtx_SpMVCSRTranspose.f90:46-145)."""
    y = np.empty(rp.size - 1)
    lib().orc_spmv_t(rp.size - 1, _p(rp), _p(ci), _p(_f64(val)), _p(_f64(x)), _p(y))
    return y


def pcg(rp, ci, val, b, x0=None, max_iter=100000, rtol=1e-10, nthreads=1):
    n = rp.size - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    ni = C.c_int32()
    rn = C.c_double()
    st = lib().orc_pcg(n, _p(rp), _p(ci), _p(val), _p(_f64(b)), _p(x), max_iter, rtol,
                       C.byref(ni), C.byref(rn), nthreads)
    return x, st, ni.value, rn.value


def pcg_mark(rp, ci, val, b, rtol, rtol_mark, max_iter=1000000, nthreads=1, report_every=0):
    """orc_pcg_mark: PCG from x = 0 to rtol; also returns the iteration and
    wall seconds at which ||r||/||b|| first fell below rtol_mark (progress
    on stderr every report_every iterations when > 0)."""
    n = rp.size - 1
    x = np.zeros(n)
    ni, rn, im, sm = C.c_int32(), C.c_double(), C.c_int32(), C.c_double()
    st = lib().orc_pcg_mark(n, _p(rp), _p(ci), _p(val), _p(_f64(b)), _p(x), max_iter, rtol,
                            rtol_mark, C.byref(ni), C.byref(rn), C.byref(im), C.byref(sm),
                            nthreads, report_every)
    return x, st, ni.value, rn.value, im.value, sm.value


def cg_transpose(rp, ci, val, rhs, tol=1e-8, max_iter=1000):
    n = rp.size - 1
    lam = np.zeros(n)
    ni = C.c_int32()
    rel = C.c_double()
    st = lib().orc_cg_transpose(n, _p(rp), _p(ci), _p(val), _p(_f64(rhs)), _p(lam), tol,
                                max_iter, C.byref(ni), C.byref(rel))
    return lam, st, ni.value, rel.value


def sensitivity(conn, coords, nu, dEdrho, u, lam, nthreads=1):
    conn, coords = _i32(conn), _f64(coords)
    g = np.empty(conn.shape[0])
    dE, uu, ll = _f64(dEdrho), _f64(u), _f64(lam)
    lib().orc_sensitivity(conn.shape[0], _p(conn), _p(coords), nu, _p(dE), _p(uu), _p(ll),
                          _p(g), nthreads)
    return g


def j2_internal_force(n_nodes, conn, coords, props, u, u_n, state_n, nthreads=1):
    """orc_j2_internal_force: (local failures, trial state, Dgp (6,6,8,ne)
    column-major per point, F_int)."""
    conn, coords = _i32(conn), _f64(coords)
    pr, uu, un, sn = j2_props(props), _f64(u), _f64(u_n), _f64(state_n)
    state = np.empty_like(sn)
    Dgp = np.empty(288 * conn.shape[0])
    Fi = np.empty(3 * n_nodes)
    nf = lib().orc_j2_internal_force(n_nodes, conn.shape[0], _p(conn), _p(coords), _p(pr),
                                     _p(uu), _p(un), _p(sn), _p(state), _p(Dgp), _p(Fi),
                                     nthreads)
    return nf, state, Dgp, Fi


# ----------------------------------------------------------------- J2
J2_PROPS_LEN = 11  # E, nu, sigma_y0, H, type, max_local_iter, K, n, eps0, Q, delta


def j2_props(props) -> np.ndarray:
    """The oracle's 11-entry props array from a shorter tuple (zero-padded:
    hardening type 0 = linear, max_local_iter 0 = 25) or from a dict of the
    ufc_ph_j2_props field names."""
    out = np.zeros(J2_PROPS_LEN)
    if isinstance(props, dict):
        keys = ["young", "nu", "sigma_y0", "hardening", "hardening_type", "max_local_iter",
                "swift_k", "swift_n", "swift_eps0", "voce_q", "voce_delta"]
        for i, k in enumerate(keys):
            out[i] = float(props.get(k, 0.0))
        return out
    p = np.asarray(props, np.float64).ravel()
    out[:p.size] = p
    return out


def j2_update(props, deps, sigma, peeq, eps_p, with_iters=False):
    """One material-point update; returns (status, sigma, peeq, eps_p, Dt,
    yielded) and, with_iters, the local Newton iterations used."""
    props, deps = j2_props(props), _f64(deps)
    sig = _f64(sigma).copy()
    pe = np.array([peeq], np.float64)
    ep = _f64(eps_p).copy()
    Dt = np.empty((6, 6))
    y = C.c_int32()
    li = C.c_int32()
    st = lib().orc_j2_update_ex(_p(props), _p(deps), _p(sig), _p(pe), _p(ep), _p(Dt),
                                C.byref(y), C.byref(li))
    if with_iters:
        return st, sig, pe[0], ep, Dt.T.copy(), y.value, li.value
    return st, sig, pe[0], ep, Dt.T.copy(), y.value


def j2_newton_step(n_nodes, conn, coords, props, bc_eq, bc_val, penalty, state_n, u,
                   max_newton=25, tol_r=1e-8, pcg_rtol=1e-12):
    conn, coords = _i32(conn), _f64(coords)
    bc_eq, bc_val = _i32(bc_eq), _f64(bc_val)
    state_n = _f64(state_n)
    state = np.empty_like(state_n)
    u = _f64(u).copy()
    nn = C.c_int32()
    rn = C.c_double()
    pr = j2_props(props)
    st = lib().orc_j2_newton_step(n_nodes, conn.shape[0], _p(conn), _p(coords), _p(pr),
                                  bc_eq.size, _p(bc_eq), _p(bc_val), penalty, _p(state_n),
                                  _p(state), _p(u), max_newton, tol_r, pcg_rtol, C.byref(nn),
                                  C.byref(rn))
    return st, u, state, nn.value, rn.value
