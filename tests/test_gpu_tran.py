"""HIP transport flux (calc_flux_tran, SRC/flux.f90:18-127) against the CPU restatement, whose
assembly is pinned in test_tran_oracle.py and whose TRANLIB property routines are pinned
bit-for-bit to the reference in test_oracle.py.  Both sides evaluate exp / log with the same code
(fortrancfd_amd/csrc/fcfd_math.h); the remaining differences are the order of the species sums of
the mixture rules (wave sums on the GPU), so the bound is 1e-12 relative to each flux component's
largest magnitude, for one evaluation and after three coupled RK3 steps."""
import numpy as np
import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu
KK = 53


@pytest.fixture(scope="module")
def mech():
    from fortrancfd_amd.lib import Mechanism
    return Mechanism()


def profile_cells(orc, n, seed):
    """(T, p, u, X) rows: a hot / cold CH4-air front with burnt products, a velocity ramp and
    seeded composition noise."""
    names = list(golden("mech")["names"])
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, n)
    Xf = np.zeros(KK)
    Xf[names.index("CH4")], Xf[names.index("O2")], Xf[names.index("N2")] = 0.095, 0.19, 0.715
    Xb = np.zeros(KK)
    Xb[names.index("CO2")], Xb[names.index("H2O")], Xb[names.index("N2")] = 0.095, 0.19, 0.715
    Xb[names.index("OH")], Xb[names.index("H")] = 0.004, 0.001
    w = 0.5 * (1 + np.tanh((x - 0.5) / 0.1))[:, None]
    X = (1 - w) * Xf + w * Xb
    X = X * (1 + 0.05 * rng.uniform(size=X.shape))
    X /= X.sum(axis=1, keepdims=True)
    T = 300 + 1900 * w[:, 0]
    cells = np.zeros((n, 3 + KK))
    cells[:, 0] = T
    cells[:, 1] = 101325.0 * (1 + 0.01 * np.sin(6 * x))
    cells[:, 2] = 20.0 * np.sin(3 * x)
    cells[:, 3:] = X
    return cells


def test_flux_tran_vs_oracle(mech, orc):
    n, lv, dx = 40, 3, 25e-6
    cells = profile_cells(orc, n + 2 * lv, 5)
    Qc = np.zeros((n + 2 * lv, 3 + KK))
    for i, (T, p, u, *X) in enumerate(cells):
        X = np.array(X)
        rho = orc.density(p, T, X)
        e = orc.specific_energy(T, X)
        Qc[i] = np.concatenate([[rho, rho * u, rho * (e + 0.5 * u * u)], rho * orc.xty(X)])
    print(f"flux_tran one block n={n}", flush=True)
    F = mech.calc_flux_tran(Qc, n, lv, dx)
    R = orc.flux_tran(Qc, n, lv, dx)
    scale = np.abs(R).max(axis=0) + 1e-300
    err = (np.abs(F - R) / scale).max()
    print(f"transport flux: max scaled difference {err:.3g}, bit-identical {np.mean(F == R):.3f}")
    assert err <= 1e-12
    assert np.abs(R[:, 2]).max() > 0 and np.abs(R[:, 3:]).max() > 0


def test_coupled_steps_with_transport_vs_oracle(mech, orc):
    from fortrancfd_amd.lib import Solver
    from oracle.pyoracle import Domain
    nblk, ncell, dx = 2, 48, 25e-6
    cells = profile_cells(orc, nblk * ncell, 6)
    kw = dict(flag_acu=2, flag_weno=2, select_riemann=2, select_time=1)
    s = Solver(mech, nblk, ncell, dx, do_tran=1, **kw)
    s.init_cells(cells)
    d = Domain(orc, cells, nblk, ncell, dx, do_tran=1, **kw)
    print("3 RK3 steps with transport, 2x48 cells", flush=True)
    for _ in range(3):
        dt_g, dt_o = s.calc_dt(0.5), d.calc_dt(0.5)
        assert abs(dt_g - dt_o) <= 1e-14 * dt_o
        s.solve_fluid(dt_o)
        d.solve_fluid(dt_o)
        d.refresh()
    Qc, qf = s.download()
    scale = np.abs(d.Qc).max(axis=(0, 1)) + 1e-300
    err = (np.abs(Qc - d.Qc) / scale).max()
    errT = (np.abs(qf[..., 2] - d.qf[..., 2]) / d.qf[..., 2]).max()
    print(f"3 steps with transport: max scaled difference Qc {err:.3g}, T {errT:.3g}")
    assert err <= 1e-12 and errT <= 1e-12
