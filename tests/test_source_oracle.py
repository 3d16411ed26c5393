"""The oracle's reaction_macks composition (fo_solve_source, restating SRC/source.f90:17-84) pinned
on the CPU by what the reference's call site fixes, independently of the integrator behind macks:

  * the integrator sees adjust_sum(Y) (:53) and its output is adjust_sum'd again (:63) before T
    and Y are written back (:65-68);
  * HRR is heat_release_rate(y_sp0, Y_new, rho, dt) with y_sp0 the raw Y before the call (:44, :70)
    and HRR_dot heat_release_rate_omegadot at the new state (:72); both functions are pinned bit for
    bit to the reference in test_oracle.py (test_hrr_golden, test_rates_golden);
  * Qc is rebuilt from qf by qfToQc (:78-82), pinned by test_eos_golden;
  * the physics of a closed adiabatic constant-volume cell: element totals, density, momentum and
    (through the energy-consistent final temperature) total energy are unchanged; a mixture of
    inert species does not move.
A dropped adjust_sum, a HRR taken from the renormalised Y0 or the wrong state, or a stale Qc fails
one of these."""
import numpy as np
import pytest

from tests.conftest import golden

KK = 53


def names():
    return list(golden("mech")["names"])


def hot_domain(orc, nblk=2, ncell=6):
    from oracle.pyoracle import Domain
    nm = names()
    cells = np.zeros((nblk * ncell, 3 + KK))
    cells[:, 0] = np.linspace(2300.0, 1400.0, nblk * ncell)
    cells[:, 1] = 101325.0 * np.linspace(1.0, 3.0, nblk * ncell)
    cells[:, 2] = np.linspace(-5.0, 5.0, nblk * ncell)
    x = np.zeros(KK)
    x[nm.index("CH4")], x[nm.index("O2")], x[nm.index("N2")] = 0.0950570342, 0.1901140684, 0.7148288973
    x[nm.index("OH")], x[nm.index("H")] = 1e-4, 1e-5
    cells[:, 3:] = x / x.sum()
    return Domain(orc, cells, nblk, ncell, 25e-6)


def interior(d, a):
    return a[:, d.lv:-d.lv, :]


def test_qc_is_qfToQc_of_the_new_qf(orc):
    d = hot_domain(orc)
    d.solve_source(1e-7)
    qf, Qc = interior(d, d.qf), interior(d, d.Qc)
    rebuilt = orc.qf_to_qc(qf.reshape(-1, qf.shape[-1])).reshape(qf.shape)
    np.testing.assert_array_equal(Qc, rebuilt)


def test_cells_are_adjust_sum_of_macks_of_adjust_sum(orc):
    """Each cell after the call equals adjust_sum(macks(dt, adjust_sum(Y0), T0, rho)), with Y0
    deliberately off the simplex (a negative entry, a sum above 1) so both adjust_sums matter."""
    d = hot_domain(orc)
    off_simplex(d)
    before = interior(d, d.qf)
    dt = 1e-7
    d.solve_source(dt)
    after = interior(d, d.qf)
    for b in range(before.shape[0]):
        for i in range(before.shape[1]):
            q0 = before[b, i]
            yT = np.concatenate([orc.adjust_sum(q0[3:]), [q0[2]]])
            out, st, rc = orc.chem_step(dt, q0[0], yT)
            assert rc == 0
            np.testing.assert_array_equal(after[b, i, 3:], orc.adjust_sum(out[:-1]))
            assert after[b, i, 2] == out[-1]
            assert after[b, i, 0] == q0[0] and after[b, i, 1] == q0[1]
    assert (after[..., 3:] >= 0).all()


def off_simplex(d):
    """Moves Y of every interior cell off the simplex: one negative entry and a sum of 1 + 1e-6."""
    qf = d._arr(d.d.qf)            # live view of the oracle's primitive state
    lv = d.lv
    qf[:, lv:-lv, 3 + names().index("HO2")] = -2e-9
    qf[:, lv:-lv, 3:] *= 1.0 + 1e-6


def test_heat_release_rates_follow_the_call_site(orc):
    """HRR from the raw Y0 (off the simplex here, so the renormalised Y0 would give another value)
    and the new state; HRR_dot at the new state."""
    d = hot_domain(orc)
    off_simplex(d)
    dt = 2e-7
    before = interior(d, d.qf)
    hrr, hrd, nf = d.solve_source(dt)
    assert nf == 0
    after = interior(d, d.qf)
    n = 0
    for b in range(before.shape[0]):
        for i in range(before.shape[1]):
            q0, q1 = before[b, i], after[b, i]
            assert hrr[n] == orc.heat_release_rate(q0[3:], q1[3:], q1[0], dt)
            assert hrr[n] != orc.heat_release_rate(orc.adjust_sum(q0[3:]), q1[3:], q1[0], dt)
            assert hrd[n] == orc.heat_release_rate_omegadot(q1[3:], q1[0], q1[2])
            n += 1
    assert np.abs(hrr).max() > 0 and np.abs(hrd).max() > 0


def test_conservation_of_a_closed_cell(orc):
    d = hot_domain(orc)
    Q0, q0 = interior(d, d.Qc), interior(d, d.qf)
    d.solve_source(5e-7)
    Q1, q1 = interior(d, d.Qc), interior(d, d.qf)
    np.testing.assert_array_equal(Q1[..., 0], Q0[..., 0])          # density
    np.testing.assert_array_equal(Q1[..., 1], Q0[..., 1])          # momentum
    np.testing.assert_allclose(Q1[..., 2], Q0[..., 2], rtol=1e-12)  # total energy
    ncf = orc.ncf()
    W = orc.weights()
    el0 = (q0[..., 3:] / W) @ ncf
    el1 = (q1[..., 3:] / W) @ ncf
    np.testing.assert_allclose(el1, el0, rtol=1e-12, atol=1e-16)
    assert np.abs(q1[..., 2] - q0[..., 2]).max() > 1.0                # the chemistry did run


def test_inert_mixture_does_not_move(orc):
    from oracle.pyoracle import Domain
    nm = names()
    cells = np.zeros((8, 3 + KK))
    cells[:, 0] = 1500.0
    cells[:, 1] = 101325.0
    cells[:, 3 + nm.index("N2")] = 0.79
    cells[:, 3 + nm.index("AR")] = 0.21
    d = Domain(orc, cells, 1, 8, 25e-6)
    q0 = interior(d, d.qf)
    Q0 = interior(d, d.Qc)
    hrr, hrd, nf = d.solve_source(1e-6)
    q1 = interior(d, d.qf)
    assert nf == 0
    np.testing.assert_array_equal(q1[..., 3:], np.array([[orc.adjust_sum(y) for y in blk] for blk in q0[..., 3:]]))
    np.testing.assert_array_equal(q1[..., 2], q0[..., 2])
    np.testing.assert_array_equal(interior(d, d.Qc), Q0)
    assert (hrr == 0).all() and (hrd == 0).all()
