"""calc_flux_tran in the CPU restatement (oracle fo_flux_tran).  The reference harness cannot run
calc_flux_tran (its viscosity comes from MCAVIS_EA, which the reference does not ship), so the
assembly of the flux is pinned here against the property routines that ARE pinned bit-for-bit to
the reference's TRANLIB (test_oracle.py::test_transport_golden) on states where the flux reduces
to one closed-form term: a uniform state has zero flux; a pure temperature jump at rest with
uniform composition leaves only the conduction term -kappa_avg dT/dx; a velocity jump at uniform
T and composition leaves tau = -4/3 mu_avg du/dx and its work tau*u_avg; a composition jump at
uniform T and rest leaves -rho_avg D_avg dY/dx per species and the enthalpy flux."""
import numpy as np

from tests.conftest import golden

KK = 53


def mixture(orc, names, **x):
    X = np.zeros(KK)
    for k, v in x.items():
        X[names.index(k)] = v
    return X / X.sum()


def cell(orc, T, p, u, X):
    Y = orc.xty(X)
    rho = orc.density(p, T, X)
    e = orc.specific_energy(T, X)
    return np.concatenate([[rho, rho * u, rho * (e + 0.5 * u * u)], rho * Y])


def two_cell_flux(orc, a, b, dx=25e-6, lv=3):
    """Flux at the a|b interface of a block of two interior cells (a, b) with lv ghosts each side
    holding copies: cells a..a (lv+1) | b..b (lv+1); interfaces 0, 1, 2 -> the middle one."""
    Qc = np.array([a] * (lv + 1) + [b] * (lv + 1))
    F = orc.flux_tran(Qc, 2, lv, dx)
    return F[1]


def props(orc, Qc):
    rho = Qc[0]
    u = Qc[1] / rho
    Y = Qc[3:] / rho
    X = orc.ytx(Y)
    T = orc.temperature_con(Qc[2] / rho - 0.5 * u * u, X)
    p = orc.pressure(rho, T, X)
    return rho, u, T, p, X, Y


def test_uniform_state_has_no_flux(orc):
    names = list(golden("mech")["names"])
    X = mixture(orc, names, CH4=0.095, O2=0.19, N2=0.715)
    c = cell(orc, 1200.0, 101325.0, 3.0, X)
    F = two_cell_flux(orc, c, c)
    assert np.all(F == 0.0)


def test_temperature_jump_is_pure_conduction(orc):
    names = list(golden("mech")["names"])
    X = mixture(orc, names, O2=0.21, N2=0.79)
    a, b = cell(orc, 300.0, 101325.0, 0.0, X), cell(orc, 2000.0, 101325.0, 0.0, X)
    dx = 25e-6
    F = two_cell_flux(orc, a, b, dx)
    _, _, Ta, _, Xa, Ya = props(orc, a)
    _, _, Tb, _, Xb, Yb = props(orc, b)
    kap = 0.5 * (orc.mcacon(Ta, Xa) * 1e-5 + orc.mcacon(Tb, Xb) * 1e-5)
    assert F[0] == 0.0 and F[1] == 0.0
    # species gradients are zero up to the round-off of Y = rho Y / rho
    np.testing.assert_allclose(F[2], -kap * (Tb - Ta) / dx, rtol=1e-9)
    assert np.abs(F[3:]).max() <= 1e-9 * abs(F[2])


def test_velocity_jump_gives_viscous_stress(orc):
    names = list(golden("mech")["names"])
    X = mixture(orc, names, O2=0.21, N2=0.79)
    a, b = cell(orc, 900.0, 101325.0, -5.0, X), cell(orc, 900.0, 101325.0, 7.0, X)
    dx = 25e-6
    F = two_cell_flux(orc, a, b, dx)
    ra, ua, Ta, _, Xa, _ = props(orc, a)
    rb, ub, Tb, _, Xb, _ = props(orc, b)
    mu = 0.5 * (orc.mcavis(Ta, Xa) * 0.1 + orc.mcavis(Tb, Xb) * 0.1)
    tau = -(4.0 / 3.0) * mu * (ub - ua) / dx
    np.testing.assert_allclose(F[1], tau, rtol=1e-12)
    np.testing.assert_allclose(F[2], tau * 0.5 * (ua + ub), rtol=1e-9)


def test_composition_jump_gives_fickian_flux(orc):
    names = list(golden("mech")["names"])
    Xa = mixture(orc, names, CH4=0.2, N2=0.8)
    Xb = mixture(orc, names, O2=0.21, N2=0.79)
    T, p, dx = 800.0, 101325.0, 25e-6
    a, b = cell(orc, T, p, 0.0, Xa), cell(orc, T, p, 0.0, Xb)
    F = two_cell_flux(orc, a, b, dx)
    ra, _, Ta, pa, Xa_, Ya = props(orc, a)
    rb, _, Tb, pb, Xb_, Yb = props(orc, b)

    def dif(p_, T_, X_, Y_):
        D = orc.mcadif(p_ * 10.0, T_, X_) * 1e-4
        with np.errstate(divide="ignore", invalid="ignore"):
            d = (1.0 - X_) / (1.0 - Y_) * D
        return np.where(Y_ == 1.0, 0.0, d)

    da = 0.5 * (dif(pa, Ta, Xa_, Ya) + dif(pb, Tb, Xb_, Yb))
    jk = -0.5 * (ra + rb) * da * (Yb - Ya) / dx
    np.testing.assert_allclose(F[3:], jk, rtol=1e-12, atol=1e-12 * np.abs(jk).max())
    hs = orc.hms(0.5 * (Ta + Tb)) * 1e-4
    np.testing.assert_allclose(F[2], (hs * jk).sum() - 0.0, rtol=1e-6)
