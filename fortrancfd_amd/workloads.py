"""Seeded synthetic workloads on the GRI-3.0 mechanism.

- coupled_cells: the coupled production domain -- nblk x ncell cells of stoichiometric CH4/air at
  1 atm and 300 K with `nhot` seeded 32-cell hot spots at T ~ U(1800, 2200) K.
- chemistry_batch: independent chemistry cells -- T ~ U(1000, 2600) K, p ~ log-U(1, 50) atm,
  dt ~ log-U(1e-8, 1e-6) s, Y a seeded mixing-line blend of fresh stoichiometric CH4/air and its
  adiabatic burnt state with Dirichlet(alpha = 50) perturbations of the minor species, renormalised.
The burnt state comes from this package's own FP64 integrator (constant volume, adiabatic, run to
equilibrium on the GPU), cached per mechanism object.
"""
from __future__ import annotations

import numpy as np

SEED = 20261202
RU_SI = 8.31451      # J/(mol K), the CHEMKIN-II RU in SI
P_ATM = 101325.0


def stoich_ch4_air_x(mech) -> np.ndarray:
    x = np.zeros(mech.kk)
    x[mech.names.index("CH4")] = 0.0950570342
    x[mech.names.index("O2")] = 0.1901140684
    x[mech.names.index("N2")] = 0.7148288973
    return x


def x_to_y(mech, x):
    w = mech.weights
    return x * w / (x * w).sum(axis=-1, keepdims=True)


def density(mech, p, T, Y):
    """Ideal-gas density [kg/m^3] from p [Pa], T [K] and mass fractions."""
    wbar = 1.0 / (Y / mech.weights).sum(axis=-1)
    return p * wbar * 1e-3 / (RU_SI * T)


def coupled_cells(mech, nblk=16, ncell=4096, nhot=3, seed=SEED):
    """Rows of (T, p, u, X...) for fcfd_init_cells."""
    n = nblk * ncell
    cells = np.zeros((n, 3 + mech.kk))
    cells[:, 0] = 300.0
    cells[:, 1] = P_ATM
    cells[:, 3:] = stoich_ch4_air_x(mech)
    rng = np.random.default_rng(seed)
    for c in rng.integers(0, n - 32, nhot):
        cells[c:c + 32, 0] = rng.uniform(1800.0, 2200.0)
    return cells


_BURNT = {}


def burnt_state(mech):
    """Adiabatic burnt state of stoichiometric CH4/air (constant volume from 1500 K, 1 atm)."""
    key = id(mech)
    if key not in _BURNT:
        y = x_to_y(mech, stoich_ch4_air_x(mech))
        T = 1500.0
        rho = density(mech, P_ATM, T, y)
        yT = np.concatenate([y, [T]])[None, :]
        for dt in (1e-4, 1e-3, 1e-2, 1e-2):
            yT, st = mech.macks(dt, rho, yT, rtol=1e-8, atol=1e-14)
        _BURNT[key] = yT[0, :-1].copy()
    return _BURNT[key]


def chemistry_batch(mech, n, seed=SEED + 3):
    """(rho [kg/m^3], dt [s], T [K], Y [n, KK]) for n independent chemistry cells."""
    rng = np.random.default_rng(seed)
    kk = mech.kk
    yf = x_to_y(mech, stoich_ch4_air_x(mech))
    yb = burnt_state(mech)
    lam = rng.uniform(0.0, 1.0, (n, 1))
    Y = lam * yf + (1.0 - lam) * yb
    major = np.zeros(kk, bool)
    for s in ("CH4", "O2", "N2", "CO2", "H2O", "CO", "H2", "AR"):
        major[mech.names.index(s)] = True
    pert = rng.dirichlet(np.full(kk, 50.0), n) * kk
    Y = np.where(major, Y, Y * pert)
    Y = np.maximum(Y, 0.0)
    Y /= Y.sum(axis=1, keepdims=True)
    T = rng.uniform(1000.0, 2600.0, n)
    p = np.exp(rng.uniform(0.0, np.log(50.0), n)) * P_ATM
    dt = np.exp(rng.uniform(np.log(1e-8), np.log(1e-6), n))
    rho = density(mech, p, T, Y)
    return rho, dt, T, Y
