"""C3D8 variants in the oracle (orc_c3d8_ke_variant), pinned by what the
mathematics fixes rather than by the formulas themselves:

- every variant: symmetric, exactly six zero-energy modes (the rigid
  motions), and the uniform-strain energy V eps:D:eps (patch test) on
  distorted elements -- REDUCED on parallelepipeds only, since its one-point
  rule (the reference's, weight 8 at the centre) integrates 8 detJ0 != V;
- REDUCED: an hourglass mode h (x) e_d carries k_hg |h|^4 = 0.1 mu V^(1/3) 64
  on the unit cube and nothing from the one-point part;
- SELECTIVE: one volumetric constraint per element (one eigenvalue grows
  with the bulk modulus as nu -> 1/2, against seven for the full element);
- INCOMPAT: pure bending of a unit cube (nu = 0) is reproduced exactly, where
  the standard element is 1.5x too stiff (shear locking);
- 27-point: equal to the 2x2x2 element wherever 2x2x2 Gauss is exact
  (parallelepipeds).
The last two tests restate the reference's own REDUCED hourglass vectors and
incompatible modes (PH_Elem_C3D8.f90:123-127, 240-259, 419-464) in numpy to
pin the defects the variants correct.
"""
import numpy as np
import pytest

XS = np.array([-1, 1, 1, -1, -1, 1, 1, -1.])
YS = np.array([-1, -1, 1, 1, -1, -1, 1, 1.])
ZS = np.array([-1, -1, -1, -1, 1, 1, 1, 1.])
UNIT = np.stack([(XS + 1) / 2, (YS + 1) / 2, (ZS + 1) / 2], 1)
PARA = UNIT @ np.array([[1.3, 0.2, 0.1], [0.0, 0.8, 0.3], [0.0, 0.0, 1.7]]).T + [0.3, -1.0, 2.0]
DIST = UNIT + 0.15 * np.random.default_rng(1).uniform(-1, 1, (8, 3))
GP = 0.577350269189626
PTS = [(a, b, c) for a in (-GP, GP) for b in (-GP, GP) for c in (-GP, GP)]
VARIANTS = [1, 2, 3, 4, 6]


def _D(E, nu):
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[range(3), range(3)] += 2 * mu
    D[3, 3] = D[4, 4] = D[5, 5] = mu
    return D


def _volume(orc, x):
    return sum(orc.dndx(x, *p)[1] for p in PTS)


def _rigid_modes(x):
    modes = []
    for d in range(3):
        u = np.zeros((8, 3))
        u[:, d] = 1.0
        modes.append(u.ravel())
    for i, j in [(0, 1), (1, 2), (0, 2)]:
        u = np.zeros((8, 3))
        u[:, i], u[:, j] = -x[:, j], x[:, i]
        modes.append(u.ravel())
    return modes


def _patch_ratio(orc, K, x, E, nu):
    A = np.random.default_rng(0).standard_normal((3, 3))
    A = A + A.T
    u = (x @ A.T).ravel()
    eps = np.array([A[0, 0], A[1, 1], A[2, 2], 2 * A[0, 1], 2 * A[1, 2], 2 * A[0, 2]])
    return (u @ K @ u) / (_volume(orc, x) * eps @ _D(E, nu) @ eps)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("shape", ["unit", "para", "dist"])
def test_variant_invariants(orc, variant, shape):
    x = {"unit": UNIT, "para": PARA, "dist": DIST}[shape]
    E, nu = 2.0e9, 0.3
    K = orc.ke_variant(x, E, nu, variant)
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-14 * scale
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-10 * ev.max()) == 6 and ev.min() > -1e-10 * ev.max()
    for m in _rigid_modes(x):
        assert abs(m @ K @ m) <= 1e-12 * scale * (m @ m)
    if variant != 2 or shape != "dist":
        assert abs(_patch_ratio(orc, K, x, E, nu) - 1.0) < 1e-12
    else:  # centre rule: 8 detJ0 in place of V
        assert abs(_patch_ratio(orc, K, x, E, nu) - 1.0) < 1e-3


def test_standard_aliases_and_unknown_variant(orc):
    K = orc.ke(DIST, 1.0, 0.25)
    for v in (0, 1, 5):
        assert np.array_equal(orc.ke_variant(DIST, 1.0, 0.25, v), K)
    with pytest.raises(ValueError):
        orc.ke_variant(DIST, 1.0, 0.25, 7)


def test_reduced_hourglass_energy_closed_form(orc):
    E, nu = 3.0, 0.25
    mu = E / (2 * (1 + nu))
    K = orc.ke_variant(UNIT, E, nu, 2)
    for h in (-XS * YS * ZS, YS * ZS, XS * ZS, XS * YS):
        for d in range(3):
            u = np.zeros((8, 3))
            u[:, d] = h
            u = u.ravel()
            assert u @ K @ u == pytest.approx(0.1 * mu * 64.0, rel=1e-13)
            # the one-point part carries none of it: no strain at the centre
            g0, _, _ = orc.dndx(UNIT, 0.0, 0.0, 0.0)
            assert np.abs(_bmat(g0) @ u).max() < 1e-15


def test_selective_has_one_volumetric_constraint(orc):
    def growing(v):
        e1 = np.linalg.eigvalsh(orc.ke_variant(DIST, 1.0, 0.4999, v))
        e2 = np.linalg.eigvalsh(orc.ke_variant(DIST, 1.0, 0.49999, v))
        return int(np.sum(e2 > 3.0 * e1 + 1e-9))
    assert growing(3) == 1
    assert growing(1) == 7


def test_incompat_pure_bending_exact(orc):
    xc = UNIT - 0.5
    kap = 1e-3
    u = np.zeros((8, 3))
    u[:, 0] = kap * xc[:, 0] * xc[:, 2]
    u[:, 2] = -kap * xc[:, 0] ** 2 / 2
    u = u.ravel()
    exact = 0.5 * 1.0 * kap ** 2 / 12.0  # E/2 kap^2 int z^2 dV, E = 1, nu = 0
    assert 0.5 * u @ orc.ke_variant(UNIT, 1.0, 0.0, 4) @ u == pytest.approx(exact, rel=1e-12)
    assert 0.5 * u @ orc.ke_variant(UNIT, 1.0, 0.0, 1) @ u == pytest.approx(1.5 * exact, rel=1e-12)


def test_gauss27_equals_2x2x2_on_parallelepipeds(orc):
    K8, K27 = orc.ke_variant(PARA, 1.0, 0.3, 1), orc.ke_variant(PARA, 1.0, 0.3, 6)
    assert np.abs(K27 - K8).max() < 1e-14 * np.abs(K8).max()
    K8, K27 = orc.ke_variant(DIST, 1.0, 0.3, 1), orc.ke_variant(DIST, 1.0, 0.3, 6)
    assert np.abs(K27 - K8).max() > 1e-6 * np.abs(K8).max()  # not exact there


def _bmat(G):
    B = np.zeros((6, 24))
    for a in range(8):
        gx, gy, gz = G[a]
        c = 3 * a
        B[0, c], B[1, c + 1], B[2, c + 2] = gx, gy, gz
        B[3, c], B[3, c + 1], B[4, c + 1], B[4, c + 2], B[5, c], B[5, c + 2] = gy, gx, gz, gy, gz, gx
    return B


def test_reference_hourglass_vectors_stiffen_rigid_rotation(orc):
    """PH_Elem_C3D8_StiffMatrixReduced with the reference's GAMMA8: its 4th
    column (1,1,1,1,-1,-1,-1,-1) is a linear field, so rigid rotations carry
    energy and xi.eta stays a zero-energy mode (7 zero eigenvalues)."""
    E, nu = 1.0, 0.3
    K = orc.ke_variant(UNIT, E, nu, 2)
    gamma8 = np.stack([-XS * YS * ZS, YS * ZS, XS * ZS, -ZS], 1)
    khg = 0.1 * E / (2 * (1 + nu))
    Kref = K.copy()
    for m in range(4):  # swap the hourglass part for the reference's
        hm = (-XS * YS * ZS, YS * ZS, XS * ZS, XS * YS)[m]
        Kref -= khg * np.kron(np.outer(hm, hm), np.eye(3))
        Kref += khg * np.kron(np.outer(gamma8[:, m], gamma8[:, m]), np.eye(3))
    ev = np.linalg.eigvalsh(Kref)
    assert np.sum(np.abs(ev) < 1e-10 * ev.max()) == 7
    rot = _rigid_modes(UNIT)[4]  # about x: u_y = -z, u_z = y
    assert rot @ Kref @ rot > 0.1 * (rot @ rot) * khg
    assert abs(rot @ K @ rot) < 1e-14


def test_reference_incompatible_modes_fail_the_patch_test(orc):
    """PH_Elem_C3D8_StiffMatrixIncompatible restated: three scalar modes
    along (1,1,1); xi(1-eta^2)(1-zeta^2) has a non-zero mean gradient, so the
    condensed unit cube is 2% soft in uniform strain."""
    E, nu = 1.0, 0.3
    D = _D(E, nu)
    Kuu, Kua, Kaa = np.zeros((24, 24)), np.zeros((24, 3)), np.zeros((3, 3))
    for xi, eta, zeta in PTS:
        g, dj, _ = orc.dndx(UNIT, xi, eta, zeta)
        B = _bmat(g)
        ox, oy, oz = 1 - xi * xi, 1 - eta * eta, 1 - zeta * zeta
        d_xi = np.array([-2 * xi * oy * oz, oy * oz, -2 * eta * xi * oz])
        d_eta = np.array([-2 * eta * ox * oz, -2 * xi * eta * oz, ox * oz])
        d_zeta = np.array([-2 * zeta * ox * oy, -2 * xi * zeta * oy, -2 * eta * zeta * ox])
        # unit cube: J = I/2, so d/dx = 2 d/dxi
        dx, dy, dz = 2 * d_xi, 2 * d_eta, 2 * d_zeta
        Be = np.stack([dx, dy, dz, dy + dx, dz + dy, dx + dz])
        Kuu += dj * B.T @ D @ B
        Kua += dj * B.T @ D @ Be
        Kaa += dj * Be.T @ D @ Be
    Kref = Kuu - Kua @ np.linalg.solve(Kaa, Kua.T)
    assert _patch_ratio(orc, Kref, UNIT, E, nu) == pytest.approx(0.979, abs=2e-3)
    assert _patch_ratio(orc, orc.ke_variant(UNIT, E, nu, 4), UNIT, E, nu) == pytest.approx(1.0, abs=1e-12)
