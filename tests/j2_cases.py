"""J2 uniaxial-stress cases with closed forms, shared by the oracle tests and
the GPU tests: the hardening laws of PH_J2_ComputeHardening
(PH_Mat_Plast_J2_Iso_Core.f90:549-605), the law's uniaxial closed form by
bisection, and a unit cube on rollers whose FE solution is homogeneous
uniaxial stress.  Props lists are the oracle's 11 entries: E, nu, sigma_y0,
H, type, max_local_iter, K, n, eps0, Q, delta."""
import math

import numpy as np


def sigma_y(props, e):
    E, nu, sy0, H, t, _, K, n, e0, Q, d = props
    if t == 2:
        return K * (e0 + e) ** n
    if t == 3:
        return sy0 + Q * (1.0 - math.exp(-d * e))
    return sy0 + H * e


def uniaxial_closed_form(props, e11):
    """Monotonic uniaxial stress: sigma = sigma_y(peeq) with
    e11 = sigma / E + peeq beyond yield, found by bisection on
    g(sigma) = sigma - sigma_y(e11 - sigma / E), which increases with sigma;
    sigma = E e11 below yield."""
    E = props[0]
    if E * e11 <= sigma_y(props, 0.0):
        return E * e11, 0.0
    lo, hi = 0.0, E * e11
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid - sigma_y(props, e11 - mid / E) > 0.0:
            hi = mid
        else:
            lo = mid
    sig = 0.5 * (lo + hi)
    return sig, e11 - sig / E


# J2 laws in MPa (E2E_C2_02_Uniaxial_J2Plastic.f90 units), so that the stiffness
# is 1e-7 of the 1e12 penalty and the prescribed values hold to that
J2_LAWS_MPA = {
    "linear": dict(young=200000.0, nu=0.3, sigma_y0=250.0, hardening=1000.0),
    "swift": dict(young=200000.0, nu=0.3, sigma_y0=500.0 * 0.01 ** 0.2, hardening_type=2,
                  swift_k=500.0, swift_n=0.2, swift_eps0=0.01),
    "voce": dict(young=200000.0, nu=0.3, sigma_y0=250.0, hardening_type=3, voce_q=150.0,
                 voce_delta=30.0),
}
UNIAXIAL_PATH = [0.001, 0.002, 0.003, 0.005, 0.01]


def uniaxial_bvp(coords):
    """Rollers on x=0, y=0, z=0 and u_z prescribed on z=1 of the unit cube: a
    homogeneous uniaxial stress state.  Returns the constrained equations, the
    value mask of the prescribed top, and the x=1 nodes."""
    x0, y0 = np.nonzero(coords[:, 0] == 0)[0], np.nonzero(coords[:, 1] == 0)[0]
    z0, z1 = np.nonzero(coords[:, 2] == 0)[0], np.nonzero(coords[:, 2] == 1)[0]
    eq = np.concatenate([3 * x0, 3 * y0 + 1, 3 * z0 + 2, 3 * z1 + 2]).astype(np.int32)
    top = np.zeros(eq.size)
    top[eq.size - z1.size:] = 1.0
    return eq, top, np.nonzero(coords[:, 0] == 1)[0]


def check_uniaxial_state(props, e33, u, state, x1):
    """Every Gauss point at the closed form of the law (independent bisection
    on sigma = sigma_y(e33 - sigma / E)): sigma_33 and peeq, the other stress
    components zero, and the lateral face at -(nu sigma / E + peeq / 2)."""
    p = [props["young"], props["nu"], props["sigma_y0"], props.get("hardening", 0.0),
         props.get("hardening_type", 1), 0, props.get("swift_k", 0.0),
         props.get("swift_n", 0.0), props.get("swift_eps0", 0.0), props.get("voce_q", 0.0),
         props.get("voce_delta", 0.0)]
    s_cf, pe_cf = uniaxial_closed_form(p, e33)
    st = state.reshape(-1, 13)
    assert np.abs(st[:, 2] - s_cf).max() < 1e-6 * s_cf, (e33, st[:, 2].min(), s_cf)
    assert np.abs(st[:, [0, 1, 3, 4, 5]]).max() < 1e-6 * s_cf
    assert np.abs(st[:, 6] - pe_cf).max() < 1e-6 * max(pe_cf, 1e-3), (e33, st[:, 6].max(), pe_cf)
    lat = -(p[1] * s_cf / p[0] + 0.5 * pe_cf)
    assert np.abs(u[3 * x1] - lat).max() < 1e-6 * abs(lat)
