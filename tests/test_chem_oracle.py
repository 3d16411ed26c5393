"""The CPU restatement of the chemistry source.  The reference's stiff integrator (MACKS in module
ODEs, called at SRC/source.f90:58) is not shipped, so its parity is unpinned; this project's own
integrator is pinned here instead: against scipy's Radau5 at rtol 1e-12 on the same right-hand side
(whose rates are bit-exact against CKWYR in test_oracle.py), element and energy conservation, the
equilibrium fixed point from the reference's Cantera fixture and the dt -> 0 limit."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

from tests.conftest import golden

KK = 53


def ch4_air(orc):
    names = list(golden("mech")["names"])
    x = np.zeros(KK)
    x[names.index("CH4")], x[names.index("O2")], x[names.index("N2")] = 0.0950570342, 0.1901140684, 0.7148288973
    return orc.xty(x), x


def rho_of(orc, x, T, p_atm):
    return p_atm * 101325.0 * (x * orc.weights()).sum() * 1e-3 / (8.31451 * T)


@pytest.mark.parametrize("T0,p_atm,dt", [(1500.0, 1.0, 1e-6), (2000.0, 1.0, 1e-7),
                                         (2400.0, 30.0, 1e-8), (1200.0, 10.0, 1e-6)])
def test_integrator_vs_radau(orc, T0, p_atm, dt):
    Y, x = ch4_air(orc)
    rho = rho_of(orc, x, T0, p_atm)
    yT = np.concatenate([Y, [T0]])
    out, st, rc = orc.chem_step(dt, rho, yT)
    assert rc == 0
    ref = solve_ivp(lambda t, y: orc.chem_rhs(rho * 1e-3, y), (0, dt), yT, method="Radau",
                    rtol=1e-12, atol=1e-20, jac=lambda t, y: orc.chem_jac(rho * 1e-3, y)).y[:, -1]
    assert abs(out[-1] - ref[-1]) <= 1e-8 * ref[-1]
    assert np.abs(out[:-1] - ref[:-1]).max() <= 1e-8


def test_conservation(orc):
    Y, x = ch4_air(orc)
    rho = rho_of(orc, x, 2200.0, 5.0)
    yT = np.concatenate([Y, [2200.0]])
    out, _, rc = orc.chem_step(5e-7, rho, yT)
    assert rc == 0
    ncf = orc.ncf()
    W = orc.weights()
    np.testing.assert_allclose((out[:-1] / W) @ ncf, (Y / W) @ ncf, rtol=1e-12, atol=1e-16)
    assert abs(out[:-1].sum() - 1.0) < 1e-12
    e0 = orc.specific_energy(2200.0, orc.adjust_sum(orc.ytx(Y)))
    ya = orc.adjust_sum(out[:-1])
    e1 = orc.specific_energy(out[-1], orc.adjust_sum(orc.ytx(ya)))
    assert abs(e1 - e0) <= 1e-11 * abs(e0)
    assert abs(out[-1] - 2200.0) > 1.0 and np.abs(out[:-1] - Y).max() > 1e-4   # the state evolved


def test_equilibrium_is_a_fixed_point(orc):
    """The reference's Cantera TP-equilibrium composition at 1300 K, 1 atm does not move."""
    g = golden("equil_tp_ch4_phi1_1atm_1300K")
    X = g["X"]
    Y = orc.xty(X)
    rho = rho_of(orc, X, 1300.0, 1.0)
    out, st, rc = orc.chem_step(1e-4, rho, np.concatenate([Y, [1300.0]]))
    assert rc == 0
    assert abs(out[-1] - 1300.0) < 1e-3
    assert np.abs(out[:-1] - Y).max() < 1e-8


def test_small_dt_limit(orc):
    """dt -> 0: dY = dt * W omega / rho to first order."""
    Y, x = ch4_air(orc)
    Y = 0.9 * Y
    names = list(golden("mech")["names"])
    Y[names.index("H")] = 0.05
    Y[names.index("OH")] = 0.05
    rho = rho_of(orc, orc.ytx(Y), 1800.0, 1.0)
    yT = np.concatenate([Y, [1800.0]])
    dt = 1e-13
    out, _, rc = orc.chem_step(dt, rho, yT, rtol=1e-10, atol=1e-20)
    f = orc.chem_rhs(rho * 1e-3, yT)
    np.testing.assert_allclose((out[:-1] - Y) / dt, f[:-1], rtol=2e-3, atol=1e-3 * np.abs(f).max())


def test_jacobian_mass_action_part(orc):
    """The analytic Jacobian matches finite differences of the right-hand side in the temperature
    column and, away from third-body / fall-off dependence, the species columns."""
    Y, x = ch4_air(orc)
    rng = np.random.default_rng(3)
    Y = 0.8 * Y + 0.2 * rng.dirichlet(np.ones(KK))
    yT = np.concatenate([Y, [1700.0]])
    rho = rho_of(orc, orc.ytx(Y), 1700.0, 2.0) * 1e-3
    J = orc.chem_jac(rho, yT)
    f0 = orc.chem_rhs(rho, yT)
    h = 1e-3
    yp = yT.copy()
    yp[-1] += h
    col = (orc.chem_rhs(rho, yp) - f0) / h
    np.testing.assert_allclose(J[:, -1], col, rtol=1e-2, atol=1e-6 * np.abs(col).max())
    # species columns: agreement of the dominant entries (frozen third-body factors aside)
    names = list(golden("mech")["names"])
    k = names.index("CH4")
    yp = yT.copy()
    dy = 1e-7
    yp[k] += dy
    colk = (orc.chem_rhs(rho, yp) - f0) / dy
    big = np.abs(colk) > 1e-2 * np.abs(colk).max()
    np.testing.assert_allclose(J[big, k], colk[big], rtol=5e-2)
