"""HIP chemistry source (one cell per wavefront, SDIRK3 + pivoted LU) against the CPU restatement
of the same integrator, and the invariants the physics fixes.  The integrator is this project's own
design (the reference's MACKS source is not shipped), so parity is GPU vs CPU restatement; the CPU
restatement is pinned to scipy's Radau5 and to the invariants in tests/test_chem_oracle.py.
Tolerances: transcendental functions (exp/log/pow) differ by ~1 ulp between the GPU's and the host's
math libraries, so bit-equality is not expected; the written bound is 1e-9 relative on T and 1e-9
absolute on Y (the integration itself runs at rtol 1e-5 / atol 1e-13)."""
import numpy as np
import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mech():
    from fortrancfd_amd.lib import Mechanism
    return Mechanism()


def mixing_line_cells(orc, n, seed):
    """Config-3-like states: T ~ U(1000, 2600), p ~ logU(1, 50) atm, dt ~ logU(1e-8, 1e-6),
    stoichiometric CH4/air partially burnt (seeded, small n)."""
    rng = np.random.default_rng(seed)
    kk = orc.kk
    names = list(golden("mech")["names"])
    x = np.zeros(kk)
    x[names.index("CH4")], x[names.index("O2")], x[names.index("N2")] = 0.0950570342, 0.1901140684, 0.7148288973
    yf = orc.xty(x)
    yb = np.zeros(kk)
    yb[names.index("CO2")], yb[names.index("H2O")], yb[names.index("N2")] = 0.1511, 0.1237, 0.7252
    lam = rng.uniform(0, 1, n)[:, None]
    Y = lam * yf + (1 - lam) * yb
    Y = np.maximum(Y * rng.dirichlet(np.full(kk, 50.0), n) * kk, 0)
    Y[:, names.index("OH")] += 1e-6 * rng.uniform(0, 1, n)
    Y /= Y.sum(axis=1, keepdims=True)
    T = rng.uniform(1000.0, 2600.0, n)
    p = np.exp(rng.uniform(0, np.log(50.0), n)) * 101325.0
    W = orc.weights()
    rho = p * (1.0 / (Y / W).sum(axis=1)) * 1e-3 / (8.31451 * T)
    dt = np.exp(rng.uniform(np.log(1e-8), np.log(1e-6), n))
    return np.column_stack([Y, T]), rho, dt


def test_macks_batch_vs_oracle(mech, orc):
    yT, rho, dt = mixing_line_cells(orc, 48, 11)
    print("macks batch: 48 cells", flush=True)
    out, st = mech.macks(dt, rho, yT)
    assert (st[:, 3] == 0).all()
    for c in range(yT.shape[0]):
        ref, rst, rc = orc.chem_step(dt[c], rho[c], yT[c])
        assert rc == 0
        assert abs(out[c, -1] - ref[-1]) <= 1e-9 * ref[-1], (c, out[c, -1], ref[-1], st[c], rst)
        assert np.abs(out[c, :-1] - ref[:-1]).max() <= 1e-9, c


def test_macks_invariants(mech, orc):
    """Element conservation, energy conservation and sum Y = 1 over 4096 cells."""
    yT, rho, dt = mixing_line_cells(orc, 4096, 12)
    out, st = mech.macks(dt, rho, yT)
    assert (st[:, 3] == 0).all() and np.isfinite(out).all()
    ncf = orc.ncf()
    W = orc.weights()
    el0 = (yT[:, :-1] / W) @ ncf
    el1 = (out[:, :-1] / W) @ ncf
    assert np.abs(el1 - el0).max() <= 1e-12 * np.abs(el0).max()
    assert np.abs(out[:, :-1].sum(axis=1) - 1.0).max() < 1e-11
    for c in range(0, 4096, 256):
        ya0 = orc.adjust_sum(yT[c, :-1])
        ya1 = orc.adjust_sum(out[c, :-1])
        e0 = orc.specific_energy(yT[c, -1], orc.adjust_sum(orc.ytx(ya0)))
        e1 = orc.specific_energy(out[c, -1], orc.adjust_sum(orc.ytx(ya1)))
        assert abs(e1 - e0) <= 1e-10 * abs(e0) + 1e-6, c


def hot_spot_cells(kk):
    """2 blocks x 32 cells of stoichiometric CH4/air at 1 atm, 300 K except 12 cells from 2200 K
    down to 1500 K (each of them needs 7 to 65 integrator attempts over dt = 2e-7 s)."""
    names = list(golden("mech")["names"])
    nblk, ncell, dx = 2, 32, 25e-6
    cells = np.zeros((nblk * ncell, 3 + kk))
    cells[:, 0] = 300.0
    cells[:12, 0] = np.linspace(2200.0, 1500.0, 12)
    cells[:, 1] = 101325.0
    x = np.zeros(kk)
    x[names.index("CH4")], x[names.index("O2")], x[names.index("N2")] = 0.0950570342, 0.1901140684, 0.7148288973
    cells[:, 3:] = x
    return cells, nblk, ncell, dx


def test_solve_source_vs_oracle(mech, orc):
    """reaction_macks on a 2-block domain with a hot spot, three calls in a row (the second and third
    start from the integrator's per-cell step-size memory on both sides): qf, Qc, HRR, HRR_dot vs
    the oracle."""
    from oracle.pyoracle import Domain
    from fortrancfd_amd.lib import Solver
    cells, nblk, ncell, dx = hot_spot_cells(mech.kk)
    d = Domain(orc, cells, nblk, ncell, dx)
    s = Solver(mech, nblk, ncell, dx, do_reaction=1)
    s.init_cells(cells)
    dt = 2e-7
    for call in range(3):
        print("solve_source call", call, flush=True)
        hrr_o, hrd_o, nf = d.solve_source(dt)
        hrr_g, hrd_g = s.solve_source(dt)
        assert nf == 0
    compare_with_oracle(s, d, hrr_g, hrd_g, hrr_o, hrd_o)


def test_solve_source_long_list_vs_oracle(mech, orc):
    """A stiff list longer than the four-wave limit, which a one-workgroup integrator grid lowers
    to 8 cells, so the 12 stiff cells go to the one-wave kernel (the mode of long lists, which the
    other domain tests never reach).  Three calls vs the oracle."""
    from oracle.pyoracle import Domain
    from fortrancfd_amd.lib import Solver
    cells, nblk, ncell, dx = hot_spot_cells(mech.kk)
    d = Domain(orc, cells, nblk, ncell, dx)
    s = Solver(mech, nblk, ncell, dx, do_reaction=1)
    s.set_chem_workgroups(1)
    s.init_cells(cells)
    dt = 2e-7
    for call in range(3):
        print("solve_source (one-workgroup grid) call", call, flush=True)
        hrr_o, hrd_o, nf = d.solve_source(dt)
        hrr_g, hrd_g = s.solve_source(dt)
        assert nf == 0
        acc, _ = s.source_accepted()
        assert (acc == 0).sum() > 8     # a long list
    compare_with_oracle(s, d, hrr_g, hrd_g, hrr_o, hrd_o)


def test_solve_source_lane_retry_vs_oracle(mech, orc):
    """Cold cells with trace radicals (1e-12 mole fraction of CO, O, C, CH3, OH, H, HO2) fail one
    explicit step of dt = 1e-8 s (error norm 2.5) but pass two half steps (0.6, then 0.3), as the
    cold region of a flame run does.  With a one-workgroup integrator grid (four-wave limit 8 cells)
    the 64-cell stiff list takes the lane-per-cell retry, which finishes 53 cells and leaves 11 to
    the implicit integrator (without a second retry); with the default grid the same list goes to
    the integrator, which retries each cell itself.  Both vs the oracle over three calls."""
    from oracle.pyoracle import Domain
    from fortrancfd_amd.lib import Solver
    names = list(golden("mech")["names"])
    for grid in (1, 0):
        cells, nblk, ncell, dx = hot_spot_cells(mech.kk)
        for sp in ("CO", "O", "C", "CH3", "OH", "H", "HO2"):
            cells[12:, 3 + names.index(sp)] = 1e-12
        cells[:, 3:] /= cells[:, 3:].sum(axis=1, keepdims=True)
        d = Domain(orc, cells, nblk, ncell, dx)
        s = Solver(mech, nblk, ncell, dx, do_reaction=1)
        s.set_chem_workgroups(grid)
        s.init_cells(cells)
        dt = 1e-8
        for call in range(3):
            print("solve_source grid", grid, "call", call, flush=True)
            hrr_o, hrd_o, nf = d.solve_source(dt)
            hrr_g, hrd_g = s.solve_source(dt)
            assert nf == 0
            if call == 0:
                acc, _ = s.source_accepted()
                st = s.chem_stats()
                assert (acc == 0).sum() == 64                           # every cell fails the screen
                assert ((st[:, 0] == 2) & (st[:, 2] == 0)).sum() == 53   # finished by the retry
        compare_with_oracle(s, d, hrr_g, hrd_g, hrr_o, hrd_o)


def compare_with_oracle(s, d, hrr_g, hrd_g, hrr_o, hrd_o):
    Qc, qf = s.download()
    lv = s.lv
    np.testing.assert_allclose(qf[:, lv:-lv, 2], d.qf[:, lv:-lv, 2], rtol=1e-9)
    np.testing.assert_allclose(qf[:, lv:-lv, 3:], d.qf[:, lv:-lv, 3:], rtol=0, atol=1e-9)
    np.testing.assert_allclose(Qc[:, lv:-lv, :], d.Qc[:, lv:-lv, :], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(hrr_g, hrr_o, rtol=1e-6, atol=1e-6 * np.abs(hrr_o).max())
    np.testing.assert_allclose(hrd_g, hrd_o, rtol=1e-6, atol=1e-6 * np.abs(hrd_o).max())


def test_coupled_advance_graph_matches_stepwise(mech):
    """Lie-split reacting steps (calc_dt, solve_fluid, solve_source, refresh) replayed from a
    captured graph equal the host-driven sequence bit-for-bit."""
    from fortrancfd_amd.lib import Solver
    kk = mech.kk
    names = list(golden("mech")["names"])
    nblk, ncell, dx = 2, 64, 25e-6
    cells = np.zeros((nblk * ncell, 3 + kk))
    cells[:, 0] = 300.0
    cells[:16, 0] = 2000.0
    cells[:, 1] = 101325.0
    cells[:, 3 + names.index("CH4")] = 0.0950570342
    cells[:, 3 + names.index("O2")] = 0.1901140684
    cells[:, 3 + names.index("N2")] = 0.7148288973
    a = Solver(mech, nblk, ncell, dx, do_reaction=1)
    b = Solver(mech, nblk, ncell, dx, do_reaction=1)
    a.init_cells(cells)
    b.init_cells(cells)
    ta = a.advance(0.5, 5)
    tb = 0.0
    for _ in range(5):
        dt = b.calc_dt(0.5)
        tb += dt
        b.solve_fluid(dt)
        b.solve_source(dt)
        b.refresh()
    Qa, qa = a.download()
    Qb, qb = b.download()
    assert ta == pytest.approx(tb, rel=1e-15)
    np.testing.assert_array_equal(Qa, Qb)
    np.testing.assert_array_equal(qa, qb)
    assert qa[:, :, 2].max() > 1990.0
