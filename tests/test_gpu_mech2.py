"""A second mechanism through the same path: Li-Dryer H2/O2 (11 species, 21 reactions; Troe fall-off,
third-body efficiencies, duplicates), read from its CHEMKIN text by fcfd_mech_load_text and run by
kernels instantiated for 11 species, against the oracle on the same mechanism's link file.  Same
bars as GRI-3.0 (tests/test_gpu_chem.py): rates of progress within 1e-14, integrator counters
identical then 1e-12, fluid + chemistry steps within 1e-12 of each field's scale."""
import numpy as np
import pytest

from tests.test_gpu_chem import REL, cell_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mechs():
    from fortrancfd_amd.lib import LIDRYER_H2, Mechanism
    from oracle.pyoracle import Oracle
    return Mechanism(LIDRYER_H2, text=True), Mechanism(LIDRYER_H2), Oracle(LIDRYER_H2)


def h2_air(m):
    x = np.zeros(m.kk)
    x[m.names.index("H2")], x[m.names.index("O2")], x[m.names.index("N2")] = 0.2958, 0.1479, 0.5563
    return x


def states(m, orc, n, seed):
    """H2/air partially burnt to water, T 900-2500 K, p 1-20 atm, dt 1e-8-1e-6 s."""
    rng = np.random.default_rng(seed)
    yf = orc.xty(h2_air(m))
    yb = np.zeros(m.kk)
    yb[m.names.index("H2O")], yb[m.names.index("N2")] = 0.2545, 0.7455
    lam = rng.uniform(0, 1, n)[:, None]
    Y = lam * yf + (1 - lam) * yb
    for sp, s in (("H", 1e-4), ("OH", 1e-3), ("O", 1e-4), ("HO2", 1e-5)):
        Y[:, m.names.index(sp)] += s * rng.uniform(0, 1, n)
    Y /= Y.sum(axis=1, keepdims=True)
    T = rng.uniform(900.0, 2500.0, n)
    p = np.exp(rng.uniform(0, np.log(20.0), n)) * 101325.0
    rho = p * (1.0 / (Y / orc.weights()).sum(axis=1)) * 1e-3 / (8.31451 * T)
    dt = np.exp(rng.uniform(np.log(1e-8), np.log(1e-6), n))
    return Y, T, rho, dt


def test_rates_vs_oracle(mechs):
    m, _, orc = mechs
    Y, T, rho, _ = states(m, orc, 64, 1)
    w, c, d, qf, qr = m.rates(rho * 1e-3, T, Y)
    for s in range(64):
        ow, oqf, oqr = orc.ckwyr(rho[s] * 1e-3, T[s], Y[s])
        rel = lambda a, b: np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
        assert rel(qf[s], oqf).max() <= 1e-14 and rel(qr[s], oqr).max() <= 1e-14, s


def test_text_and_link_file_give_identical_results(mechs):
    mt, mb, orc = mechs
    Y, T, rho, dt = states(mt, orc, 32, 2)
    yT = np.column_stack([Y, T])
    a, sa = mt.macks(dt, rho, yT)
    b, sb = mb.macks(dt, rho, yT)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(sa, sb)


def test_macks_vs_oracle(mechs):
    m, _, orc = mechs
    Y, T, rho, dt = states(m, orc, 96, 3)
    yT = np.column_stack([Y, T])
    out, st = m.macks(dt, rho, yT)
    refs, rst = [], []
    for c in range(96):
        r, s_, rc = orc.chem_step(dt[c], rho[c], yT[c])
        assert rc == 0
        refs.append(r)
        rst.append((s_["steps"], s_["rejects"], s_["newton"]))
    same = (st[:, :3] == np.array(rst)).all(axis=1)
    err = cell_errors(out, np.array(refs))
    print(f"H2/O2: identical counters in {same.sum()} of 96 cells; max difference {err.max():.3g}")
    assert same.all() and err.max() <= REL
    assert (st[:, 2] > 0).sum() > 10          # the implicit integrator was exercised


@pytest.mark.parametrize("strang,do_tran", [(False, 0), (True, 1)])
def test_reacting_steps_vs_oracle(mechs, strang, do_tran):
    """Three production steps of a 2 x 32-cell H2/air domain with a 1400-1800 K hot spot, Lie
    without transport and Strang with transport, against the oracle's driver steps."""
    from fortrancfd_amd.lib import Solver
    from oracle.pyoracle import Domain
    from tests.test_gpu_configs import compare_fields
    m, _, orc = mechs
    cells = np.zeros((64, 3 + m.kk))
    cells[:, 0] = 300.0
    cells[:12, 0] = np.linspace(1800.0, 1400.0, 12)
    cells[:, 1] = 101325.0
    cells[:, 3:] = h2_air(m)
    kw = dict(flag_acu=2, flag_weno=2, select_riemann=2, select_time=1)
    s = Solver(m, 2, 32, 25e-6, do_reaction=1, do_tran=do_tran, **kw)
    s.set_step(strang=strang, hrr_each_step=True)
    s.init_cells(cells)
    d = Domain(orc, cells, 2, 32, 25e-6, do_tran=do_tran, **kw)
    dts, t = [], 0.0
    for _ in range(3):
        t1 = s.advance(0.5, 1, t)
        dts.append((t1 - t, d.step(0.5, react=True, strang=strang)))
        t = t1
    dts = np.array(dts)
    np.testing.assert_allclose(dts[:, 0], dts[:, 1], rtol=1e-13)
    compare_fields(s, d, dts[:, 1], dts[:, 1], orc)
    acc, _ = s.source_accepted()
    assert (acc != 2).sum() > 0
