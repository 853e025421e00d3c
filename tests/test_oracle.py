"""Pins the FP64 CPU oracle (oracle/) before it is trusted as the checker.

Each test fixes a fact that a plausible slip in the restatement would break:
reference golden vectors (CSR pattern / values / SpMV from the reference's
own NM_Mtx_Sparse.f90 build), closed forms, invariants and brute force.
Reference paths are relative to its root.
"""
import math
import os

import numpy as np
import pytest

from j2_cases import (J2_LAWS_MPA, UNIAXIAL_PATH, check_uniaxial_state, sigma_y as _sigma_y,
                      uniaxial_bvp, uniaxial_closed_form as _uniaxial_closed_form)

UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
GP3 = [(-math.sqrt(0.6), 5 / 9), (0.0, 8 / 9), (math.sqrt(0.6), 5 / 9)]


def distorted(seed=1, amp=0.2):
    rng = np.random.default_rng(seed)
    return UNIT * np.array([1.3, 0.8, 1.1]) + amp * rng.uniform(-1, 1, (8, 3))


# ------------------------------------------------------------------ shape functions
def test_shape_partition_of_unity_and_nodal_delta(orc):
    rng = np.random.default_rng(0)
    for xi, eta, zeta in rng.uniform(-1, 1, (20, 3)):
        N, dN = orc.shape(xi, eta, zeta)
        assert abs(N.sum() - 1.0) < 1e-15          # QUARTER would give 2
        assert np.abs(dN.sum(axis=0)).max() < 1e-15
    # N_a(node b) = delta_ab (Abaqus node order, PH_Elem_C3D8.f90:1930-1933)
    nat = 2 * UNIT - 1
    for b in range(8):
        N, _ = orc.shape(*nat[b])
        assert np.allclose(N, np.eye(8)[b], atol=0, rtol=0)


def test_unit_cube_centroid_gradients(orc):
    # HARNESS_Elem_C3D8_PatchTest.f90:259-266: dN/dx = +-0.25 at the centroid
    g, detJ, ok = orc.dndx(UNIT, 0.0, 0.0, 0.0)
    assert ok and abs(detJ - 0.125) < 1e-15
    assert np.allclose(np.abs(g), 0.25, rtol=0, atol=1e-15)
    assert np.allclose(g, 0.25 * (2 * UNIT - 1), rtol=0, atol=1e-15)


def test_dndx_reproduces_linear_field_on_distorted_element(orc):
    # for u = A x, sum_a dN_a/dx_j u_a,i = A_ij at any point: fails if J^-1
    # were applied where J^-T is needed
    xe = distorted(3, 0.25)
    A = np.array([[1.0, 2.0, -0.5], [0.3, -1.0, 0.7], [2.0, 0.1, 0.4]])
    u = xe @ A.T
    for xi, eta, zeta in [(0.1, -0.3, 0.5), (-0.7, 0.2, 0.9), (0.577, 0.577, -0.577)]:
        g, _, ok = orc.dndx(xe, xi, eta, zeta)
        assert ok
        grad = u.T @ g  # (3,3): sum_a u_a,i g_a,j
        assert np.allclose(grad, A, rtol=0, atol=1e-12)


def test_volume_integral(orc):
    xe = distorted(5, 0.15)
    # 2x2x2 sum of detJ (weights 1) equals the trilinear volume, exactly
    # integrated; compare with the 3x3x3 rule
    v2 = sum(orc.dndx(xe, sx * 0.577350269189626, sy * 0.577350269189626,
                      sz * 0.577350269189626)[1]
             for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1))
    v3 = sum(wa * wb * wc * orc.dndx(xe, a, b, c)[1]
             for a, wa in GP3 for b, wb in GP3 for c, wc in GP3)
    assert abs(v2 - v3) < 1e-13 * abs(v3)


# ------------------------------------------------------------------ element stiffness
def _ke_gauss3(xe, E, nu, orc):
    """Independent Bᵀ D B with 3x3x3 Gauss (exact for boxes like 2x2x2)."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[np.arange(3, 6), np.arange(3, 6)] = mu
    K = np.zeros((24, 24))
    for a, wa in GP3:
        for b, wb in GP3:
            for c, wc in GP3:
                g, dj, _ = orc.dndx(xe, a, b, c)
                B = np.zeros((6, 24))
                for n in range(8):
                    gx, gy, gz = g[n]
                    B[0, 3 * n], B[1, 3 * n + 1], B[2, 3 * n + 2] = gx, gy, gz
                    B[3, 3 * n], B[3, 3 * n + 1] = gy, gx
                    B[4, 3 * n + 1], B[4, 3 * n + 2] = gz, gy
                    B[5, 3 * n], B[5, 3 * n + 2] = gz, gx
                K += wa * wb * wc * dj * B.T @ D @ B
    return K


def test_ke_closed_form_unit_cube(orc):
    # E = 1, nu = 0 unit cube: K(1x,1x) = int (dN1/dx)^2 + (1/2)((dN1/dy)^2 + (dN1/dz)^2)
    # = 1/9 + (1/2)(2/9) = 2/9 with N1 = (1-x)(1-y)(1-z)
    K = orc.ke(UNIT, 1.0, 0.0)
    assert abs(K[0, 0] - 2.0 / 9.0) < 1e-15
    # the same at nu = 0.3, E = 210e9 against an independent 3x3x3 integration
    K = orc.ke(UNIT * np.array([2.0, 0.5, 1.5]), 210e9, 0.3)
    K3 = _ke_gauss3(UNIT * np.array([2.0, 0.5, 1.5]), 210e9, 0.3, orc)
    assert np.abs(K - K3).max() < 1e-12 * np.abs(K3).max()


def test_ke_symmetry_rigid_modes_psd(orc):
    xe = distorted(7, 0.2)
    K = orc.ke(xe, 210e9, 0.3)
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() < 1e-14 * scale
    # six rigid-body modes: 3 translations, 3 rotations
    modes = []
    for d in range(3):
        t = np.zeros((8, 3))
        t[:, d] = 1
        modes.append(t.ravel())
    for axis in range(3):
        w = np.zeros(3)
        w[axis] = 1
        modes.append(np.cross(w, xe).ravel())
    for m in modes:
        assert np.abs(K @ m).max() < 1e-12 * scale * np.abs(m).max()
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-10 * ev[-1]) == 6 and ev.min() > -1e-10 * ev[-1]


def test_ke_patch_uniform_strain(orc):
    # HARNESS_Elem_C3D8_PatchTest.f90 TC-3 analogue: uniform strain on a
    # distorted element -> Ke u = int B^T sigma with constant sigma
    xe = distorted(11, 0.2)
    E, nu = 210e9, 0.3
    eps11 = 1e-3
    u = np.zeros((8, 3))
    u[:, 0] = eps11 * xe[:, 0]
    K = orc.ke(xe, E, nu)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    sig = np.array([lam + 2 * mu, lam, lam, 0, 0, 0]) * eps11
    f = orc.fint(xe, np.tile(sig, (8, 1)))
    assert np.abs(K @ u.ravel() - f).max() < 1e-12 * np.abs(f).max()
    # and the strain is recovered at every Gauss point
    eps = orc.strain(xe, u.ravel())
    assert np.allclose(eps, np.tile([eps11, 0, 0, 0, 0, 0], (8, 1)), rtol=0, atol=1e-15)


# ------------------------------------------------------------------ CSR vs reference golden
@pytest.mark.parametrize("mesh", ["1x1x1", "2x1x1", "3x2x2", "20x4x4"])
def test_pattern_matches_reference_build(orc, golden_dir, mesh):
    """NM_CSR_AssembleFromElements from the reference's own NM_Mtx_Sparse.f90
    (built by oracle/ref.mk): same row_ptr / col_idx, same assembled values for
    integer element matrices, same SpMV."""
    g = np.load(os.path.join(golden_dir, f"ref_csr_{mesh}.npz"))
    nx, ny, nz = (int(v) for v in g["mesh"])
    coords, conn = orc.mesh_box(nx, ny, nz)
    rp, ci = orc.csr_pattern(coords.shape[0], conn)
    assert np.array_equal(rp, g["ia"] - 1)
    assert np.array_equal(ci, g["ja"] - 1)
    # literal triplet merge with the same integer Ke(i,j) = 1 + (7e+3i+j) mod 11
    ne = conn.shape[0]
    e1 = np.arange(1, ne + 1)[:, None, None]
    i1 = np.arange(1, 25)[None, :, None]
    j1 = np.arange(1, 25)[None, None, :]
    Ke = 1.0 + (7 * e1 + 3 * i1 + j1) % 11
    dofs = (3 * conn[:, :, None] + np.arange(3)[None, None, :]).reshape(ne, 24)
    ti = np.repeat(dofs[:, :, None], 24, axis=2).ravel()
    tj = np.repeat(dofs[:, None, :], 24, axis=1).ravel()
    rp2, ci2, val = orc.csr_from_triplets(3 * coords.shape[0], ti, tj, Ke.ravel())
    assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci)
    assert np.array_equal(val, g["a"])
    x = 1.0 + (5 * np.arange(1, rp.size)) % 13
    assert np.array_equal(orc.spmv(rp, ci, val, x), g["y"])


def test_pattern_nnz_closed_form(orc):
    # 27-point node stencil: nnz = 9 * prod_i (3 (n_i + 1) - 2)
    for nx, ny, nz in [(4, 3, 2), (20, 4, 4), (7, 7, 7)]:
        coords, conn = orc.mesh_box(nx, ny, nz)
        rp, ci = orc.csr_pattern(coords.shape[0], conn)
        assert ci.size == 9 * (3 * nx + 1) * (3 * ny + 1) * (3 * nz + 1)
        assert all(np.all(np.diff(ci[rp[r]:rp[r + 1]]) > 0) for r in range(0, rp.size - 1, 7))


def test_assembly_equals_triplet_merge(orc):
    coords, conn = orc.mesh_box(3, 2, 2, 1.5, 1.0, 1.0)
    rng = np.random.default_rng(2)
    E = rng.uniform(1e9, 2e9, conn.shape[0])
    rp, ci = orc.csr_pattern(coords.shape[0], conn)
    val = orc.assemble(conn, coords, E, 0.3, rp, ci)
    K = np.zeros((3 * coords.shape[0],) * 2)
    for e in range(conn.shape[0]):
        ke = orc.ke(coords[conn[e]], E[e], 0.3)
        d = (3 * conn[e][:, None] + np.arange(3)).ravel()
        K[np.ix_(d, d)] += ke
    dense = np.zeros_like(K)
    for r in range(rp.size - 1):
        dense[r, ci[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    assert np.abs(dense - K).max() < 1e-13 * np.abs(K).max()
    # OpenMP atomic variant: same values up to summation order
    val4 = orc.assemble(conn, coords, E, 0.3, rp, ci, nthreads=4)
    assert np.abs(val4 - val).max() < 1e-13 * np.abs(val).max()


# ------------------------------------------------------------------ loads / BC
def _cantilever(orc):
    coords, conn = orc.mesh_box(20, 4, 4, 10.0, 2.0, 2.0)
    elem = np.arange(conn.shape[0]).reshape(4, 4, 20)[:, :, 19].ravel()
    return coords, conn, elem


def test_loads_total_and_lumping(orc):
    coords, conn, elem = _cantilever(orc)
    F = np.zeros(3 * coords.shape[0])
    orc.face_loads_traction(elem, np.full(elem.size, 5), conn, coords, [0, 0, -250.0], F)
    assert abs(F.sum() + 1e3) < 1e-12 and abs(F[2::3].sum() + 1e3) < 1e-12
    # 16 faces of area 0.25: 15.625 per face-node contribution
    tip = np.nonzero(np.abs(coords[:, 0] - 10) < 1e-12)[0]
    corner = tip[(coords[tip, 1] == 0) & (coords[tip, 2] == 0)][0]
    assert abs(F[3 * corner + 2] + 15.625) < 1e-12
    # consistent pressure equals area/4 lumping on flat faces; +p along +x here
    F2 = np.zeros_like(F)
    orc.face_loads_pressure(elem, np.full(elem.size, 5), conn, coords, 250.0, F2)
    F3 = np.zeros_like(F)
    orc.face_loads_traction(elem, np.full(elem.size, 5), conn, coords, [250.0, 0, 0], F3)
    assert np.abs(F2 - F3).max() < 1e-12


def test_penalty_closed_form(orc):
    # one element, all 24 dofs pinned by penalty P at value v: K_pen u = P v
    # -> u = (K + P I)^-1 P v, checked against the dense solve
    coords, conn = orc.mesh_box(1, 1, 1)
    rp, ci = orc.csr_pattern(8, conn)
    val = orc.assemble(conn, coords, np.array([1e9]), 0.3, rp, ci)
    F = np.zeros(24)
    v = np.linspace(-1, 1, 24) * 1e-3
    orc.apply_penalty(np.arange(24), v, 1e12, rp, ci, val, F)
    ke = orc.ke(coords[conn[0]], 1e9, 0.3)
    d = (3 * conn[0][:, None] + np.arange(3)).ravel()
    K = np.zeros((24, 24))
    K[np.ix_(d, d)] = ke
    K += 1e12 * np.eye(24)
    assert np.allclose(F, 1e12 * v, rtol=0, atol=0)
    x, st, it, rn = orc.pcg(rp, ci, val, F, rtol=1e-14)
    assert st == 0 and np.abs(x - np.linalg.solve(K, F)).max() < 1e-12 * np.abs(v).max()


# ------------------------------------------------------------------ PCG
def test_pcg_against_dense_and_stopping_rule(orc):
    coords, conn, elem = _cantilever(orc)
    n = 3 * coords.shape[0]
    rp, ci = orc.csr_pattern(coords.shape[0], conn)
    val = orc.assemble(conn, coords, np.full(conn.shape[0], 210e9), 0.3, rp, ci)
    clamp = np.nonzero(coords[:, 0] == 0)[0]
    eq = (3 * clamp[:, None] + np.arange(3)).ravel()
    F = np.zeros(n)
    orc.face_loads_traction(elem, np.full(elem.size, 5), conn, coords, [0, 0, -250.0], F)
    orc.apply_penalty(eq, np.zeros(eq.size), 1e12, rp, ci, val, F)
    K = np.zeros((n, n))
    for r in range(n):
        K[r, ci[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    u_ref = np.linalg.solve(K, F)
    x, st, it, rn = orc.pcg(rp, ci, val, F, rtol=1e-10)
    assert st == 0 and rn < 1e-10
    r_true = np.linalg.norm(F - K @ x) / np.linalg.norm(F)
    assert r_true < 2e-10
    assert np.linalg.norm(x - u_ref) / np.linalg.norm(u_ref) < 1e-6
    # first k with ||r||/||b|| < rtol: one iteration less does not converge
    x2, st2, it2, rn2 = orc.pcg(rp, ci, val, F, rtol=1e-10, max_iter=it - 1)
    assert st2 == 1 and it2 == it - 1 and rn2 >= 1e-10
    # Euler-Bernoulli tip deflection P L^3 / (3 E I) within 5% (E2E_C2_01 check)
    tip = np.nonzero(coords[:, 0] == 10)[0]
    w = x[3 * tip + 2].mean()
    w_eb = -1e3 * 10 ** 3 / (3 * 210e9 * (2 * 2 ** 3 / 12))
    assert abs(w - w_eb) / abs(w_eb) < 0.05


def test_uniaxial_unit_cube(orc):
    # tests/ufc_core/test_c3d8_uniaxial.f90: 100 MPa on a unit cube,
    # u_z = sigma/E = 5e-4 (E = 200e9), rollers on the three symmetry faces
    coords, conn = orc.mesh_box(2, 2, 2)
    n = 3 * coords.shape[0]
    rp, ci = orc.csr_pattern(coords.shape[0], conn)
    val = orc.assemble(conn, coords, np.full(8, 200e9), 0.3, rp, ci)
    eq = np.concatenate([3 * np.nonzero(coords[:, d] == 0)[0] + d for d in range(3)])
    top = np.arange(8).reshape(2, 2, 2)[1].ravel()
    F = np.zeros(n)
    orc.face_loads_pressure(top, np.full(4, 1), conn, coords, 100e6, F)
    orc.apply_penalty(eq, np.zeros(eq.size), 1e20, rp, ci, val, F)
    x, st, it, rn = orc.pcg(rp, ci, val, F, rtol=1e-14)
    uz_top = x[3 * np.nonzero(coords[:, 2] == 1)[0] + 2]
    assert np.allclose(uz_top, 5e-4, rtol=1e-8, atol=0)
    ux_side = x[3 * np.nonzero(coords[:, 0] == 1)[0]]
    assert np.allclose(ux_side, -0.3 * 5e-4, rtol=1e-8, atol=0)


def test_cg_transpose_matches_pcg(orc):
    # NM_CG_Solve_Transpose on symmetric K: lambda = K^-1 rhs; for compliance
    # (rhs = f) lambda = u
    coords, conn = orc.mesh_box(4, 3, 3)
    rp, ci = orc.csr_pattern(coords.shape[0], conn)
    val = orc.assemble(conn, coords, np.full(conn.shape[0], 1e9), 0.3, rp, ci)
    eq = (3 * np.nonzero(coords[:, 2] == 0)[0][:, None] + np.arange(3)).ravel()
    F = np.random.default_rng(4).normal(size=3 * coords.shape[0])
    orc.apply_penalty(eq, np.zeros(eq.size), 1e12, rp, ci, val, F)
    u, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-14)
    lam, st2, it2, rel = orc.cg_transpose(rp, ci, val, F, tol=1e-13, max_iter=100000)
    assert st == 0 and st2 == 0 and rel < 1e-13
    assert np.linalg.norm(lam - u) / np.linalg.norm(u) < 1e-9


# ------------------------------------------------------------------ sensitivity
def test_sensitivity_matches_finite_differences(orc):
    # compliance c = f.u, E_e = 1e-3 + rho^3 1e9: dc/drho_e = -3 rho^2 1e9 u_e^T K_e0 u_e
    coords, conn = orc.mesh_box(3, 3, 3)
    n = 3 * coords.shape[0]
    rho = orc.field_uniform(27, 7, 0.3, 1.0)
    rp, ci = orc.csr_pattern(coords.shape[0], conn)
    eq = (3 * np.nonzero(coords[:, 2] == 0)[0][:, None] + np.arange(3)).ravel()
    top = np.arange(27).reshape(3, 3, 3)[2].ravel()

    def compliance(r):
        E = 1e-3 + r ** 3 * 1e9
        val = orc.assemble(conn, coords, E, 0.3, rp, ci)
        F = np.zeros(n)
        orc.face_loads_pressure(top, np.full(9, 1), conn, coords, 1e6, F)
        f0 = F.copy()
        orc.apply_penalty(eq, np.zeros(eq.size), 1e12, rp, ci, val, F)
        u, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-15, max_iter=20000)
        return f0 @ u, u

    c0, u = compliance(rho)
    g = orc.sensitivity(conn, coords, 0.3, 3 * rho ** 2 * 1e9, u, u)
    for e in (0, 13, 26):
        h = 1e-4
        rp_, rm_ = rho.copy(), rho.copy()
        rp_[e] += h
        rm_[e] -= h
        fd = (compliance(rp_)[0] - compliance(rm_)[0]) / (2 * h)
        assert abs(fd - g[e]) < 1e-5 * abs(g[e])


# ------------------------------------------------------------------ J2
def test_j2_elastic_uniaxial_strain(orc):
    # tests/test_UMAT_J2.f90:78: sigma_11 = 282.69230769230769 (E=210000, nu=0.3, eps=1e-3)
    st, sig, pe, ep, Dt, y = orc.j2_update([210000.0, 0.3, 1e9, 0.0], [1e-3, 0, 0, 0, 0, 0],
                                           np.zeros(6), 0.0, np.zeros(6))
    assert st == 0 and y == 0
    assert abs(sig[0] - 282.69230769230769) < 1e-9


def test_j2_return_to_yield_surface(orc):
    # HARNESS_Mat_Plastic_VonMises.f90:54-68,143: q = sigma_y after return (H = 0)
    st, sig, pe, ep, Dt, y = orc.j2_update([200e9, 0.3, 250e6, 0.0], [2e-3, 0, 0, 0, 0, 0],
                                           np.zeros(6), 0.0, np.zeros(6))
    s = sig.copy()
    s[:3] -= sig[:3].mean()
    q = math.sqrt(1.5 * (s[0] ** 2 + s[1] ** 2 + s[2] ** 2 + 2 * (s[3] ** 2 + s[4] ** 2 +
                                                                  s[5] ** 2)))
    assert st == 0 and y == 1 and abs(q - 250e6) / 250e6 < 1e-10


def _uniaxial_stress(orc, props, eps11_path):
    """Drives eps11 with sigma22 = sigma33 = 0 by Newton on the lateral strains
    (E2E_C2_02_Uniaxial_J2Plastic.f90 loading)."""
    sig, pe, ep = np.zeros(6), 0.0, np.zeros(6)
    eps = np.zeros(6)
    out = []
    for e11 in eps11_path:
        target = eps.copy()
        target[0] = e11
        lat = target[1:3].copy()
        for _ in range(50):
            d = target.copy()
            d[1:3] = lat
            st, s1, p1, e1, Dt, y = orc.j2_update(props, d - eps, sig, pe, ep)
            r = s1[1:3]
            if np.abs(r).max() < 1e-12 * max(1.0, abs(s1[0])):
                break
            lat -= np.linalg.solve(Dt[1:3, 1:3], r)
        eps = d
        sig, pe, ep = s1, p1, e1
        out.append((sig.copy(), pe))
    return out


def test_j2_uniaxial_stress_closed_form(orc):
    # E2E_C2_02: E=200000, nu=0.3, sigma_y=250, H=1000, eps = .001 .002 .003 .005.
    # Uniaxial stress with linear hardening: sigma = (sigma_y + H eps)/(1 + H/E)
    # beyond yield, E eps below it (monotonic loading).
    E, sy, H = 200000.0, 250.0, 1000.0
    path = [0.001, 0.002, 0.003, 0.005]
    res = _uniaxial_stress(orc, [E, 0.3, sy, H], path)
    for e11, (sig, pe) in zip(path, res):
        if e11 <= sy / E:
            assert abs(sig[0] - E * e11) < 1e-9 * E * e11 and pe < 1e-12
        else:
            expect = (sy + H * e11) / (1 + H / E)
            assert abs(sig[0] - expect) < 1e-9 * expect
            assert abs(pe - (e11 - expect / E)) < 1e-9


def test_j2_consistent_tangent_by_finite_differences(orc):
    props = [200e9, 0.3, 250e6, 2e9]
    sig0 = np.array([180e6, -40e6, 30e6, 60e6, -25e6, 10e6])
    pe0, ep0 = 0.002, np.zeros(6)
    deps = np.array([1.5e-3, -4e-4, 2e-4, 9e-4, -3e-4, 5e-4])
    st, sig, pe, ep, Dt, y = orc.j2_update(props, deps, sig0, pe0, ep0)
    assert st == 0 and y == 1
    h = 1e-9
    fd = np.zeros((6, 6))
    for j in range(6):
        dp, dm = deps.copy(), deps.copy()
        dp[j] += h
        dm[j] -= h
        sp = orc.j2_update(props, dp, sig0, pe0, ep0)[1]
        sm = orc.j2_update(props, dm, sig0, pe0, ep0)[1]
        fd[:, j] = (sp - sm) / (2 * h)
    assert np.abs(fd - Dt).max() < 1e-5 * np.abs(Dt).max()
    assert np.abs(Dt - Dt.T).max() < 1e-6 * np.abs(Dt).max()


# hardening laws of PH_J2_ComputeHardening (PH_Mat_Plast_J2_Iso_Core.f90:549-605)
# as 11-entry oracle props: E, nu, sigma_y0, H, type, max_local_iter, K, n,
# eps0, sigma_inf, delta
J2_LAWS = {
    "linear": [200e9, 0.3, 250e6, 2e9, 1, 0, 0, 0, 0, 0, 0],
    "swift": [200e9, 0.3, 500e6 * 0.01 ** 0.2, 0, 2, 0, 500e6, 0.2, 0.01, 0, 0],
    "voce": [200e9, 0.3, 250e6, 0, 3, 0, 0, 0, 0, 150e6, 30.0],
}


@pytest.mark.parametrize("law", ["linear", "swift", "voce"])
def test_j2_hardening_uniaxial_closed_form(orc, law):
    """Each hardening law under uniaxial stress against the closed form of
    the law itself (independent bisection, not the radial return): beyond
    yield sigma_11 = sigma_y(peeq) and eps_11 = sigma_11 / E + peeq."""
    props = J2_LAWS[law]
    path = [0.0005, 0.002, 0.004, 0.008, 0.02, 0.05]
    res = _uniaxial_stress(orc, props, path)
    for e11, (sig, pe) in zip(path, res):
        s_cf, pe_cf = _uniaxial_closed_form(props, e11)
        assert abs(sig[0] - s_cf) < 1e-9 * s_cf, (law, e11, sig[0], s_cf)
        assert abs(pe - pe_cf) < 1e-9 * max(pe_cf, 1e-3), (law, e11, pe, pe_cf)
        assert np.abs(sig[1:]).max() < 1e-6 * s_cf
    # the laws really differ past yield
    assert res[-1][0][0] > 1.05 * _sigma_y(props, 0.0)


@pytest.mark.parametrize("law", ["linear", "swift", "voce"])
def test_j2_hardening_consistent_tangent_fd(orc, law):
    """The consistent tangent with H'(peeq_n + dg) against central
    differences of the stress update, for each law, from a plastic state."""
    props = J2_LAWS[law]
    sig0 = np.array([180e6, -40e6, 30e6, 60e6, -25e6, 10e6])
    pe0, ep0 = 0.01, np.zeros(6)
    deps = np.array([1.5e-3, -4e-4, 2e-4, 9e-4, -3e-4, 5e-4])
    st, sig, pe, ep, Dt, y, li = orc.j2_update(props, deps, sig0, pe0, ep0, with_iters=True)
    assert st == 0 and y == 1
    h = 1e-9
    fd = np.zeros((6, 6))
    for j in range(6):
        dp, dm = deps.copy(), deps.copy()
        dp[j] += h
        dm[j] -= h
        sp = orc.j2_update(props, dp, sig0, pe0, ep0)[1]
        sm = orc.j2_update(props, dm, sig0, pe0, ep0)[1]
        fd[:, j] = (sp - sm) / (2 * h)
    assert np.abs(fd - Dt).max() < 1e-5 * np.abs(Dt).max(), law
    assert np.abs(Dt - Dt.T).max() < 1e-6 * np.abs(Dt).max()
    # a tangent with H' at peeq_n instead of peeq_n + dg fails the same check
    if law != "linear":
        G = props[0] / (2 * (1 + props[1]))
        s = sig.copy()
        s[:3] -= s[:3].mean()
        nrm = math.sqrt(s[0] ** 2 + s[1] ** 2 + s[2] ** 2 + 2 * (s[3] ** 2 + s[4] ** 2 +
                                                                 s[5] ** 2))
        nv = s / nrm
        def htan(e):
            return (_sigma_y(props, e + 1e-7) - _sigma_y(props, e - 1e-7)) / 2e-7
        th2_new = 3 * G / (3 * G + htan(pe))
        th2_old = 3 * G / (3 * G + htan(pe0))
        wrong = Dt + 2 * G * (th2_new - th2_old) * np.outer(nv, nv)
        assert np.abs(fd - wrong).max() > 1e-4 * np.abs(Dt).max()


def test_j2_local_newton_iterations_and_limit(orc):
    """Linear hardening returns in one local iteration (the first guess is
    exact); Voce and Swift iterate; a local limit below what a step needs is
    a failure (status 2, the reference's pnewdt cut-back signal,
    PH_Mat_Plast_J2_Iso_Core.f90:419-427), and the same limit passes a
    step small enough."""
    deps = np.array([4e-3, -1.2e-3, -1.2e-3, 0, 0, 0])
    z = np.zeros(6)
    for law, lo in (("linear", 1), ("swift", 3), ("voce", 3)):
        st, sig, pe, ep, Dt, y, li = orc.j2_update(J2_LAWS[law], deps, z, 0.0, z,
                                                   with_iters=True)
        assert st == 0 and y == 1
        assert (li == 1) if law == "linear" else (li >= lo), (law, li)
    limited = list(J2_LAWS["voce"])
    limited[5] = 2
    st, *_ = orc.j2_update(limited, deps, z, 0.0, z)
    assert st == 2
    # a step just past yield converges within two local iterations
    small = np.array([1.4e-3, -0.42e-3, -0.42e-3, 0, 0, 0])
    st, sig, pe, ep, Dt, y, li = orc.j2_update(limited, small, z, 0.0, z, with_iters=True)
    assert st == 0 and y == 1 and li <= 2


def test_j2_newton_converges_quadratically(orc):
    coords, conn = orc.mesh_box(2, 2, 2, jitter=0.1, seed=5)
    bottom = np.nonzero(coords[:, 2] == 0)[0]
    top = np.nonzero(coords[:, 2] == 1)[0]
    eq = np.concatenate([(3 * bottom[:, None] + np.arange(3)).ravel(), 3 * top + 2])
    props = [200e9, 0.3, 250e6, 2e9]
    state = np.zeros(8 * 8 * 13)
    u = np.zeros(3 * coords.shape[0])
    its = []
    for inc in range(1, 4):
        val = np.concatenate([np.zeros(3 * bottom.size), np.full(top.size, 1e-3 * inc)])
        st, u, state_new, nn, rn = orc.j2_newton_step(coords.shape[0], conn, coords, props, eq,
                                                      val, 1e12, state, u)
        assert st == 0
        its.append(nn)
        state = state_new
    assert max(its) <= 6
    peeq = state.reshape(-1, 13)[:, 6]
    assert peeq.max() > 0  # yielded by 0.3% strain (yield strain 0.125%)


@pytest.mark.parametrize("law", ["linear", "swift", "voce"])
def test_j2_newton_uniaxial_stress_closed_form(orc, law):
    """orc_j2_newton_step on a boundary-value problem with a closed form: a
    2x2x2 unit cube on rollers, u_z on top driven through 0.1-1% strain
    (E2E_C2_02 loading).  The FE solution is homogeneous uniaxial stress, so
    each increment's Gauss-point stresses, plastic strains and lateral
    displacement equal the law's own closed form.  Newton stops at 1e-7 of
    the internal force: the penalty rows resolve P ulp(u) = 1e12 x 1.7e-18,
    about 1e-8 of it here."""
    props = J2_LAWS_MPA[law]
    coords, conn = orc.mesh_box(2, 2, 2)
    eq, top, x1 = uniaxial_bvp(coords)
    state = np.zeros(8 * 8 * 13)
    u = np.zeros(3 * coords.shape[0])
    for e33 in UNIAXIAL_PATH:
        st, u, state, nn, rn = orc.j2_newton_step(coords.shape[0], conn, coords, props, eq,
                                                  e33 * top, 1e12, state, u, tol_r=1e-7)
        assert st == 0, (law, e33)
        check_uniaxial_state(props, e33, u, state, x1)
    assert state.reshape(-1, 13)[:, 6].min() > 0.0  # every point yielded


def test_spmv_transpose_against_dense(orc):
    """orc_spmv_t against the dense A^T x on a random non-symmetric CSR, and
    equal to orc_spmv on a symmetric stiffness matrix."""
    rng = np.random.default_rng(4)
    n = 40
    dense = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.2)
    rows, cols = np.nonzero(dense)
    rp, ci, val = orc.csr_from_triplets(n, rows, cols, dense[rows, cols])
    x = rng.standard_normal(n)
    assert np.allclose(orc.spmv_t(rp, ci, val, x), dense.T @ x, rtol=0, atol=1e-13)
    coords, conn = orc.mesh_box(3, 2, 2)
    rp, ci = orc.csr_pattern(coords.shape[0], conn)
    K = orc.assemble(conn, coords, np.full(conn.shape[0], 1e9), 0.3, rp, ci)
    x = rng.standard_normal(rp.size - 1)
    assert np.abs(orc.spmv_t(rp, ci, K, x) - orc.spmv(rp, ci, K, x)).max() < 1e-12 * 1e9


@pytest.mark.parametrize("mesh", ["1x1x1", "2x1x1", "3x2x2", "20x4x4"])
def test_spmv_transpose_matches_reference_transpose(orc, golden_dir, mesh):
    """orc_spmv_t against the reference's own NM_CSR_Transpose applied to the
    golden (non-symmetric, integer) matrix: exact in FP64."""
    import os
    g = np.load(os.path.join(golden_dir, f"ref_csr_{mesh}.npz"))
    n = g["ia"].size - 1
    x = 1.0 + (5 * np.arange(1, n + 1)) % 13
    yt = orc.spmv_t(g["ia"] - 1, g["ja"] - 1, g["a"], x)
    assert np.array_equal(yt, g["yt"])
    assert not np.array_equal(g["yt"], g["y"])


def test_reference_sparse_kernels_cantilever_match_oracle(orc):
    """oracle/_ref/ref_cantilever (the reference's own COO element assembly,
    NM_COO_to_CSR and NM_CSR_MatVec_Optimized, compiled from the reference
    tree; element matrices, loads, penalty and the PCG loop restated in
    Fortran) solves the 20x4x4 cantilever to 1e-10 in the oracle's number of
    iterations and to its displacements (1e-10)."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                       "_ref", "ref_cantilever")
    if not os.path.exists(exe):
        pytest.skip("reference-built program not present (needs the reference tree to build)")
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "u.txt")
        subprocess.run([exe, "1", out], check=True, timeout=300)
        txt = open(out).read().split()
    ndof, nnz, iters, rep = (int(v) for v in txt[:4])
    u = np.array([float(v) for v in txt[4 + 2 * rep:]])
    coords, conn = orc.mesh_box(20, 4, 4, 10.0, 2.0, 2.0)
    rp, ci = orc.csr_pattern(coords.shape[0], conn)
    val = orc.assemble(conn, coords, np.full(conn.shape[0], 210e9), 0.3, rp, ci)
    assert ndof == 3 * coords.shape[0] and nnz == val.size
    F = np.zeros(ndof)
    tip = np.arange(19, 320, 20)  # elements at i = nx - 1
    orc.face_loads_traction(tip.astype(np.int32), np.full(tip.size, 5, np.int32), conn, coords,
                            (0.0, 0.0, -1e3 / 4.0), F)
    x0 = np.nonzero(coords[:, 0] == 0.0)[0]
    eq = (3 * x0[:, None] + np.arange(3)).ravel().astype(np.int32)
    orc.apply_penalty(eq, np.zeros(eq.size), 1e12, rp, ci, val, F)
    uo, st, it, _ = orc.pcg(rp, ci, val, F, rtol=1e-10)
    assert st == 0 and it == iters
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-10


def test_degenerate_points_skipped_as_reference(orc):
    """PH_Elem_C3D8_StiffMatrix skips a Gauss point with |det J| <= 1e-12
    (PH_Elem_C3D8.f90:775): a zero-height element has K_e = 0; an element
    collapsed to a wedge (two top nodes on the bottom) keeps the points
    with a nonzero det J; the assembly reports degenerate elements and
    skips their points only when asked to."""
    flat = UNIT.copy()
    flat[4:, 2] = 0.0
    assert np.abs(orc.ke(flat, 1e9, 0.3)).max() == 0.0
    wedge = UNIT.copy()
    wedge[[5, 6], 2] = 0.0    # the x = 1 top edge on the bottom face
    Kw = orc.ke(wedge, 1e9, 0.3)
    # the points whose det J stays above 1e-12 contribute: independent sum
    Kr = np.zeros((24, 24))
    g = 1 / math.sqrt(3)
    for xi in (-g, g):
        for eta in (-g, g):
            for zeta in (-g, g):
                G, dj, ok = orc.dndx(wedge, xi, eta, zeta)
                if not ok or abs(dj) <= 1e-12:
                    continue
                B = np.zeros((6, 24))
                for n in range(8):
                    gx, gy, gz = G[n]
                    B[0, 3 * n], B[1, 3 * n + 1], B[2, 3 * n + 2] = gx, gy, gz
                    B[3, 3 * n], B[3, 3 * n + 1] = gy, gx
                    B[4, 3 * n + 1], B[4, 3 * n + 2] = gz, gy
                    B[5, 3 * n], B[5, 3 * n + 2] = gz, gx
                lam, mu = 1e9 * 0.3 / (1.3 * 0.4), 1e9 / 2.6
                D = np.zeros((6, 6))
                D[:3, :3] = lam
                D[np.arange(3), np.arange(3)] += 2 * mu
                D[np.arange(3, 6), np.arange(3, 6)] = mu
                Kr += dj * B.T @ D @ B
    assert np.abs(Kw - Kr).max() <= 1e-12 * np.abs(Kr).max()
    coords = np.concatenate([UNIT, UNIT + [0, 0, 1.0]])
    coords[4:8, 2] = 0.0      # the first element's top face on its bottom
    conn = np.array([[0, 1, 2, 3, 4, 5, 6, 7], [4, 5, 6, 7, 12, 13, 14, 15]], np.int32)
    coords[12:, 2] = 1.0
    rp, ci = orc.csr_pattern(16, conn)
    with pytest.raises(RuntimeError):
        orc.assemble(conn, coords, np.full(2, 1e9), 0.3, rp, ci)
    val = orc.assemble(conn, coords, np.full(2, 1e9), 0.3, rp, ci, allow_degenerate=True)
    assert np.abs(val).max() > 0.0
