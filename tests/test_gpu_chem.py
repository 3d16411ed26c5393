"""HIP chemistry source against the CPU restatement of the same integrator (the reference's MACKS
source is not shipped, so parity is GPU vs CPU restatement; the restatement is pinned to scipy's
Radau5 and to the invariants in tests/test_chem_oracle.py), and the device rate code against the
reference's own CKWYR / CKCDYR outputs (tests/golden/rates.npz).

Both sides evaluate exp / log / log10 / pow with the same code (fortrancfd_amd/csrc/fcfd_math.h), so
the remaining differences come from the order of floating-point sums (wave-wide tree sums on the
GPU, sequential loops on the CPU), i.e. round-off.  Comparisons of the integrator first require the
per-cell counters (steps, rejects, Newton iterations) to be identical, then hold the result to 1e-12
relative: T relative, each mass fraction relative to max(|Y_k|, 1e-13) (the integrator's absolute
tolerance; below it a mass fraction is noise at the integrator's own tolerance).  Heat-release
rates are differences of nearly equal sums; they are held to 1e-12 of the magnitude of their terms.
"""
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


REL = 1e-12
Y_FLOOR = 1e-13


def cell_errors(out, ref):
    """Per cell, the larger of the T relative difference and max_k |dY_k| / max(|Y_k|, Y_FLOOR);
    prints where the largest one sits."""
    et = np.abs(out[..., -1] - ref[..., -1]) / np.abs(ref[..., -1])
    eyk = np.abs(out[..., :-1] - ref[..., :-1]) / np.maximum(np.abs(ref[..., :-1]), Y_FLOOR)
    ey = eyk.max(axis=-1)
    c = int(np.argmax(np.maximum(et, ey)))
    k = int(np.argmax(eyk[c]))
    print(f"  largest difference: cell {c}, T {et[c]:.3g}, species {k} {ey[c]:.3g} "
          f"(Y {ref[c, k]:.3g} vs {out[c, k]:.3g})", flush=True)
    return np.maximum(et, ey)


def test_macks_batch_vs_oracle(mech, orc):
    yT, rho, dt = mixing_line_cells(orc, 96, 11)
    print("macks batch: 96 cells", flush=True)
    out, st = mech.macks(dt, rho, yT)
    assert (st[:, 3] == 0).all()
    refs, rsts = [], []
    for c in range(yT.shape[0]):
        ref, rst, rc = orc.chem_step(dt[c], rho[c], yT[c])
        assert rc == 0
        refs.append(ref)
        rsts.append((rst["steps"], rst["rejects"], rst["newton"]))
    refs, rsts = np.array(refs), np.array(rsts)
    same = (st[:, :3] == rsts).all(axis=1)
    err = cell_errors(out, refs)
    print(f"identical counters in {same.sum()} of {same.size} cells; max difference {err.max():.3g}")
    assert same.all(), np.nonzero(~same)
    assert err.max() <= REL, err.max()


def test_rates_vs_reference_golden(mech, orc_libm):
    """Device CKWYR / CKCDYR (the integrator's rate code) against the reference's own outputs for
    40 states (tests/golden/rates.npz): rates of progress within 1e-12 relative; production,
    creation and destruction rates within 1e-12 of the magnitude of the terms they sum (the wave
    adds them in another order than CKWYR's loop, and they cancel)."""
    g = golden("rates")
    ii, kk = mech.ii, mech.kk
    rho = g["rho"] * 1e-3
    w, c, d, qf, qr = mech.rates(rho, g["T"], g["Y"])
    par, nu, nk, ns, nr = orc_libm.arrhenius()
    inc = np.zeros((kk, ii))      # |nu| of each species in each reaction (both sides counted)
    for i in range(ii):
        for n in range(6):
            if nk[i, n] > 0:
                inc[nk[i, n] - 1, i] += abs(nu[i, n])
    worst = 0.0
    for s_ in range(len(g["T"])):
        ref = g["out"][s_]
        rqf, rqr = ref[3 * kk:3 * kk + ii], ref[3 * kk + ii:3 * kk + 2 * ii]
        rel = lambda a, b: np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
        assert rel(qf[s_], rqf).max() <= REL and rel(qr[s_], rqr).max() <= REL, s_
        scale = inc @ (np.abs(rqf) + np.abs(rqr))
        for got, want in ((w[s_], ref[:kk]), (c[s_], ref[kk:2 * kk]), (d[s_], ref[2 * kk:3 * kk])):
            e = np.abs(got - want) / np.maximum(scale, 1e-300)
            assert e.max() <= REL, s_
            worst = max(worst, e.max())
    print("device rates vs reference CKWYR: worst scaled difference", worst)


def test_rates_vs_oracle(mech, orc):
    """Device rates against the restatement with the same math: every rate of progress of a
    reaction without third body agrees bit for bit; third-body and fall-off rows use the total
    concentration, which the wave sums in another order (1e-14 relative)."""
    g = golden("rates")
    kk = mech.kk
    rho = g["rho"] * 1e-3
    w, c, d, qf, qr = mech.rates(rho, g["T"], g["Y"])
    par, nu, nk, ns, nr = orc.arrhenius()
    exact = 0
    for s_ in range(len(g["T"])):
        ow, oqf, oqr = orc.ckwyr(rho[s_], g["T"][s_], g["Y"][s_])
        rel = np.maximum(np.abs(qf[s_] - oqf) / np.maximum(np.abs(oqf), 1e-300),
                         np.abs(qr[s_] - oqr) / np.maximum(np.abs(oqr), 1e-300))
        assert rel.max() <= 1e-14, s_
        exact += int((rel == 0).sum())
    print(f"rates of progress bit-identical: {exact} of {len(g['T']) * mech.ii}")
    assert exact >= 0.8 * len(g["T"]) * mech.ii


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
    compare_with_oracle(s, d, hrr_g, hrd_g, hrr_o, hrd_o, orc, dt)


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
    compare_with_oracle(s, d, hrr_g, hrd_g, hrr_o, hrd_o, orc, dt)


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
                assert (acc != 2).sum() == 64                           # every cell fails the screen
                assert ((st[:, 0] == 2) & (st[:, 2] == 0)).sum() == 53   # finished by the retry
                assert (acc == 3).sum() == 53 and (acc == 0).sum() == 11
        compare_with_oracle(s, d, hrr_g, hrd_g, hrr_o, hrd_o, orc, dt)


def compare_with_oracle(s, d, hrr_g, hrd_g, hrr_o, hrd_o, orc=None, dt=None):
    """qf / Qc / HRR / HRR_dot of the last solve_source: counters first, then 1e-12 (see the module
    docstring for the scales)."""
    Qc, qf = s.download()
    lv = s.lv
    st_g, st_o = s.chem_stats(), d.chem_stats()
    same = (st_g[:, :3] == st_o[:, :3]).all(axis=1)
    print(f"identical counters in {same.sum()} of {same.size} cells", flush=True)
    assert same.all(), (np.nonzero(~same), st_g[~same], st_o[~same])
    g = qf[:, lv:-lv, 2:].reshape(-1, qf.shape[-1] - 2)
    o = d.qf[:, lv:-lv, 2:].reshape(-1, qf.shape[-1] - 2)
    yt_g = np.concatenate([g[:, 1:], g[:, :1]], axis=1)
    yt_o = np.concatenate([o[:, 1:], o[:, :1]], axis=1)
    err = cell_errors(yt_g, yt_o)
    print(f"max difference of T, Y: {err.max():.3g}", flush=True)
    assert err.max() <= REL
    rho_max = np.abs(d.Qc[:, lv:-lv, 0]).max()
    np.testing.assert_allclose(Qc[:, lv:-lv, :], d.Qc[:, lv:-lv, :], rtol=REL, atol=REL * Y_FLOOR * rho_max)
    if orc is not None:
        eh, ed = hrr_errors(orc, d, dt, hrr_g, hrd_g, hrr_o, hrd_o)
        print(f"HRR, HRR_dot: max scaled difference {eh:.3g}, {ed:.3g}", flush=True)
        assert eh <= REL and ed <= REL


def hrr_errors(orc, d, dt, hrr_g, hrd_g, hrr_o, hrd_o, field=False):
    """Largest HRR and HRR_dot differences, each relative to the magnitude of the terms it sums:
    HRR = -sum rho h_k (Y_k - Y0_k) / dt against sum |rho h_k Y_k| / dt (the terms before the
    difference); HRR_dot = -sum W_k h_k (c_k - d_k) against sum |W_k h_k| (c_k + d_k).  With
    field=True (runs where the fluid has carried trace species below the integrator's absolute
    tolerance into the cold mixture, whose HRR_dot they set and whose relative differences are
    not resolved), each is relative to the domain's largest term magnitude instead."""
    lv = d.lv
    h = orc.hms(298.15)
    W = orc.weights()
    rho = d.qf[:, lv:-lv, 0].reshape(-1)
    T = d.qf[:, lv:-lv, 2].reshape(-1)
    Y = d.qf[:, lv:-lv, 3:].reshape(-1, W.size)
    sc_h = (np.abs(h) * Y).sum(axis=1) * rho * 1e-3 / dt * 0.1
    sc_d = np.zeros_like(sc_h)
    for c in range(Y.shape[0]):
        cd, dd = orc.ckcdyr(rho[c] * 1e-3, T[c], Y[c])
        sc_d[c] = (np.abs(W * h) * (cd + dd)).sum() * 0.1
    # a cell whose terms are 100 orders of magnitude below the domain's largest (the cold mixture at
    # 300 K, where rates of progress are subnormal) is judged against that floor
    if field:
        sc_h = np.full_like(sc_h, sc_h.max())
        sc_d = np.full_like(sc_d, sc_d.max())
    else:
        sc_h = np.maximum(sc_h, 1e-100 * sc_h.max())
        sc_d = np.maximum(sc_d, 1e-100 * sc_d.max())
    eh = np.abs(hrr_g - hrr_o) / sc_h
    ed = np.abs(hrd_g - hrd_o) / sc_d
    c = int(np.argmax(ed))
    print(f"  largest HRR_dot difference: cell {c}, {hrd_g[c]:.6g} vs {hrd_o[c]:.6g}, scale {sc_d[c]:.3g}",
          flush=True)
    return float(eh.max()), float(ed.max())


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
