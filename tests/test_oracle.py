"""The CPU oracle against the reference's own outputs (golden fixtures written by
tests/golden/make_golden.py from the amdflang-built reference) and against what the mathematics
fixes.  No GPU needed."""
import numpy as np
import pytest

from tests.conftest import golden

KK = 53


def maxrel(a, b, floor=1e-300):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float((np.abs(a - b) / np.maximum(np.abs(b), floor)).max())


# ---------------------------------------------------------------------------------------------
# golden vectors (reference Fortran, bit-for-bit)
# ---------------------------------------------------------------------------------------------
def test_mechanism_tables_match_ckinit(orc):
    g = golden("mech")
    iv, dv = g["iv"], g["dv"]
    ck = iv[:85]
    leni, lenr = iv[85], iv[86]
    ick = iv[87:87 + leni]
    rck = dv[:lenr]
    kk, ii, mxsp, mxtp, ncp2, npar = ck[1], ck[2], ck[3], ck[5], ck[8], ck[10]
    assert (kk, ii) == (orc.kk, orc.ii) == (53, 325)
    nc_wt, nc_tt, nc_aa, nc_co = ck[57] - 1, ck[58] - 1, ck[59] - 1, ck[60] - 1
    ic_nu, ic_nk, ic_ns, ic_nr = ck[34] - 1, ck[35] - 1, ck[36] - 1, ck[37] - 1
    np.testing.assert_array_equal(orc.weights(), rck[nc_wt:nc_wt + kk])
    tt, aa = orc.nasa()
    np.testing.assert_array_equal(tt.ravel(), rck[nc_tt:nc_tt + kk * mxtp])
    np.testing.assert_array_equal(aa.ravel(), rck[nc_aa:nc_aa + kk * (mxtp - 1) * ncp2])
    par, nu, nk, ns, nr = orc.arrhenius()
    np.testing.assert_array_equal(par.ravel(), rck[nc_co:nc_co + ii * (npar + 1)])
    np.testing.assert_array_equal(nu.ravel(), ick[ic_nu:ic_nu + ii * mxsp])
    np.testing.assert_array_equal(nk.ravel(), ick[ic_nk:ic_nk + ii * mxsp])
    np.testing.assert_array_equal(ns, ick[ic_ns:ic_ns + ii])
    np.testing.assert_array_equal(nr, ick[ic_nr:ic_nr + ii])


def test_reconstruction_golden(orc):
    g = golden("recon")
    np.testing.assert_array_equal(orc.recon(g["s"]), g["out"])


@pytest.mark.parametrize("variant", ["libm", "fm"])
def test_thermo_golden(orc, orc_libm, variant):
    orc = orc_libm if variant == "libm" else orc
    g = golden("thermo")
    for i, (T, rho, Y) in enumerate(zip(g["T"], g["rho"], g["Y"])):
        X = orc.ytx(Y)
        Xa = orc.adjust_sum(X)
        Ya = orc.adjust_sum(Y)
        p = orc.pressure(rho, T, Xa)
        e = orc.specific_energy(T, Xa)
        mine = np.concatenate([X, Xa, Ya, [p, e, orc.cp(T, Xa), orc.cv(T, Xa),
                                           orc.sonic_speed(rho, p, T, Xa),
                                           orc.temperature_con(e, Xa)],
                               orc.hms(T), orc.smh(T), orc.cpml(T), orc.uml(T)])
        assert maxrel(mine, g["out"][i], 1e-30) <= 1e-15, (i, T)


def test_eos_golden(orc):
    g = golden("eos")
    np.testing.assert_array_equal(orc.qf_to_qc(g["qf"]), g["qf2Qc"])
    np.testing.assert_array_equal(orc.qc_to_qf(g["Qc"]), g["Qc2qf"])


def test_rates_golden(orc_libm):
    """The restatement built with the C library's math reproduces the reference's CKWYR, CKCDYR
    and HRR_dot bit for bit."""
    orc = orc_libm
    g = golden("rates")
    ii = orc.ii
    for i in range(len(g["T"])):
        rho, T, Y = g["rho"][i], g["T"][i], g["Y"][i]
        w, qf, qr = orc.ckwyr(rho * 1e-3, T, Y)
        c, d = orc.ckcdyr(rho * 1e-3, T, Y)
        ref = g["out"][i]
        np.testing.assert_array_equal(w, ref[:KK])
        np.testing.assert_array_equal(c, ref[KK:2 * KK])
        np.testing.assert_array_equal(d, ref[2 * KK:3 * KK])
        np.testing.assert_array_equal(qf, ref[3 * KK:3 * KK + ii])
        np.testing.assert_array_equal(qr, ref[3 * KK + ii:3 * KK + 2 * ii])
        assert orc.heat_release_rate_omegadot(Y, rho, T) == ref[-1]


@pytest.mark.parametrize("variant", ["libm", "fm"])
def test_hrr_golden(orc, orc_libm, variant):
    orc = orc_libm if variant == "libm" else orc
    g = golden("hrr")
    for r, h in zip(g["rows"], g["hrr"]):
        assert orc.heat_release_rate(r[2:2 + KK], r[2 + KK:], r[0], r[1]) == h


def test_transport_golden(orc_libm):
    orc = orc_libm
    g = golden("trans")
    for i in range(len(g["T"])):
        T, p, X = g["T"][i], g["p"][i], g["X"][i]
        mine = [orc.mcavis(T, X) * 0.1, orc.mcacon(T, X) * 1e-5, *(orc.mcadif(p * 10, T, X) * 1e-4)]
        np.testing.assert_array_equal(mine, g["out"][i])


def _terms_scale(nu_rows, q):
    """Per species, the sum of |nu * q| over the terms CKWYR adds (the scale of its cancellation)."""
    return nu_rows @ np.abs(q)


def test_rates_golden_shared_math(orc):
    """The restatement with the library's own exp/log/log10/pow (the build the GPU is compared
    with) against the reference's CKWYR / CKCDYR: rates of progress within 1e-13 relative, and the
    production rates within 1e-13 of the magnitude of the terms they sum (they cancel)."""
    g = golden("rates")
    ii = orc.ii
    par, nu, nk, ns, nr = orc.arrhenius()
    inc = np.zeros((KK, ii))
    for i in range(ii):
        for n in range(6):
            if nk[i, n] > 0:
                inc[nk[i, n] - 1, i] += abs(nu[i, n])
    worst = 0.0
    for i in range(len(g["T"])):
        rho, T, Y = g["rho"][i], g["T"][i], g["Y"][i]
        w, qf, qr = orc.ckwyr(rho * 1e-3, T, Y)
        c, d = orc.ckcdyr(rho * 1e-3, T, Y)
        ref = g["out"][i]
        rqf, rqr = ref[3 * KK:3 * KK + ii], ref[3 * KK + ii:3 * KK + 2 * ii]
        assert maxrel(qf, rqf) <= 1e-13 and maxrel(qr, rqr) <= 1e-13, i
        scale = _terms_scale(inc, rqf) + _terms_scale(inc, rqr)
        err = np.abs(w - ref[:KK]) / np.maximum(scale, 1e-300)
        assert err.max() <= 1e-13, i
        assert maxrel(c, ref[KK:2 * KK]) <= 1e-13 and maxrel(d, ref[2 * KK:3 * KK]) <= 1e-13, i
        worst = max(worst, err.max(), maxrel(qf, rqf))
    print("shared-math rates vs reference: worst relative difference", worst)


def test_transport_golden_shared_math(orc):
    g = golden("trans")
    for i in range(len(g["T"])):
        T, p, X = g["T"][i], g["p"][i], g["X"][i]
        mine = [orc.mcavis(T, X) * 0.1, orc.mcacon(T, X) * 1e-5, *(orc.mcadif(p * 10, T, X) * 1e-4)]
        assert maxrel(mine, g["out"][i]) <= 1e-13, i


def test_flux_golden(orc):
    g = golden("flux")
    for ci, (acu, wen, mus, riem) in enumerate(g["cases"]):
        F, G = orc.flux_advection(g["qf"], int(g["n"]), int(g["lv"]), float(g["dx"]),
                                  flag_acu=int(acu), flag_weno=int(wen), flag_muscl=int(mus),
                                  select_riemann=int(riem))
        np.testing.assert_array_equal(F, g["Fplus"][ci])
        np.testing.assert_array_equal(G, g["godunov"][ci])


RUNS = ["sod_weno5z_4blk", "sod_muscl_100blk", "sod_euler_hll", "sod_rk4_lxf", "ch4_hot_weno5"]


def run_oracle(orc, g, nsteps=None):
    from oracle.pyoracle import Domain
    nblk, ncell, ns, lv, acu, wen, mus, riem, tim, cflf = [int(x) for x in g["params"]]
    d = Domain(orc, g["cells"], nblk, ncell, float(g["dx"]), flag_acu=acu, flag_weno=wen,
               flag_muscl=mus, select_riemann=riem, select_time=tim)
    init = (d.Qc, d.qf)
    dts = []
    for _ in range(nsteps or ns):
        dt = d.calc_dt(float(g["cfl"])) if cflf else float(g["dt"])
        dts.append(dt)
        d.solve_fluid(dt)
        d.refresh()
    return d, init, np.array(dts)


@pytest.mark.parametrize("name", RUNS)
def test_runs_golden(orc, name):
    g = golden("run_" + name)
    d, (Qc0, qf0), dts = run_oracle(orc, g)
    np.testing.assert_array_equal(Qc0, g["Qc0"])
    np.testing.assert_array_equal(qf0, g["qf0"])
    np.testing.assert_array_equal(dts, g["dts"])
    np.testing.assert_array_equal(d.Qc, g["Qc"])
    np.testing.assert_array_equal(d.qf, g["qf"])


def test_sod_config_golden_first_steps(orc):
    """The 1000-cell SOD configuration: 25 of the fixture's 200 steps keep the CPU suite quick; the
    dt sequence must agree step for step."""
    g = golden("run_sod_weno5js_hllc")
    d, _, dts = run_oracle(orc, g, nsteps=25)
    np.testing.assert_array_equal(dts, g["dts"][:25])


# ---------------------------------------------------------------------------------------------
# what the mathematics fixes
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("variant", ["libm", "fm"])
def test_equilibrium_detailed_balance(orc, orc_libm, variant):
    """At the reference's Cantera TP-equilibrium state (1300 K, 1 atm) every reversible reaction's
    net rate vanishes relative to its forward rate: pins K_c (NASA Gibbs energies, pressure
    factor, sign conventions) independently of CKLIB."""
    orc = orc_libm if variant == "libm" else orc
    g = golden("equil_tp_ch4_phi1_1atm_1300K")
    X = g["X"]
    Y = orc.xty(X)
    rho = 1.01325e6 * (X * orc.weights()).sum() / (8.31451e7 * 1300.0)
    _, qf, qr = orc.ckwyr(rho, 1300.0, Y)
    _, _, _, ns, _ = orc.arrhenius()
    rev = ns > 0
    live = rev & (qf > 0)
    assert rev.sum() == 309 and live.sum() >= 300
    assert (np.abs(qf - qr)[live] / qf[live]).max() < 1e-11
    assert np.all(qr[rev & (qf == 0)] == 0)


def test_weno5_reproduces_polynomials(orc):
    rng = np.random.default_rng(0)
    for _ in range(20):
        a, b = rng.standard_normal(2)
        s = a + b * np.arange(8.0)
        out = orc.recon(s[None, :])[0]
        # interface between s[2] and s[3]: linear data -> a + b*2.5 from both sides (JS and Z)
        for col in (2, 3, 4, 5):
            assert abs(out[col] - (a + b * 2.5)) < 1e-12 * (1 + abs(a) + abs(b))
    s = np.full(8, 3.25)
    assert np.all(orc.recon(s[None, :])[0][:16] == 3.25)


def test_weno5_smooth_limit_is_fifth_order_stencil(orc):
    """On smooth data the JS weights tend to the linear weights (1,6,3)/10, i.e. the upwind
    five-point stencil (2,-13,47,27,-3)/60."""
    x = np.arange(6) * 1e-3
    s = np.sin(x) * 1e-4 + 1.0
    out = orc.recon(np.concatenate([s, [0.0, 0.0]])[None, :])[0]
    lin = (2 * s[0] - 13 * s[1] + 47 * s[2] + 27 * s[3] - 3 * s[4]) / 60.0
    assert abs(out[2] - lin) < 1e-12


def test_hllc_consistency(orc):
    """F(U, U) equals the physical flux (uniform stencil -> no jump at the interface)."""
    g = golden("flux")
    row = g["qf"][5].copy()
    qf = np.repeat(row[None, :], 10, axis=0)
    F, _ = orc.flux_advection(qf, 2, 4, 1e-3, flag_acu=2, flag_weno=2, select_riemann=2)
    Qc = orc.qf_to_qc(row[None, :])[0]
    X = orc.adjust_sum(orc.ytx(row[3:]))
    p = orc.pressure(row[0], row[2], X)
    phys = np.concatenate([[row[0] * row[1], row[0] * row[1] ** 2 + p, (Qc[2] + p) * row[1]],
                           row[0] * row[3:] * row[1]])
    np.testing.assert_allclose(F[1], phys, rtol=1e-13, atol=1e-300)


def test_wall_run_conserves_mass_and_species(orc):
    """Reflective walls: total mass and every species' mass are conserved to round-off."""
    g = golden("run_ch4_hot_weno5")
    d, (Qc0, _), _ = run_oracle(orc, g, nsteps=10)
    lv = d.lv
    tot0 = Qc0[:, lv:-lv, :].sum(axis=(0, 1))
    tot1 = d.Qc[:, lv:-lv, :].sum(axis=(0, 1))
    for n in [0] + list(range(3, 3 + KK)):
        assert abs(tot1[n] - tot0[n]) <= 1e-12 * max(abs(tot0[n]), 1e-30) + 1e-300
    assert abs(tot1[2] - tot0[2]) <= 1e-12 * abs(tot0[2])


def test_temperature_newton_round_trip(orc):
    rng = np.random.default_rng(3)
    for T in [250.0, 600.0, 1000.0, 1500.0, 2600.0, 3400.0]:
        y = rng.dirichlet(np.ones(KK))
        X = orc.adjust_sum(orc.ytx(y))
        e = orc.specific_energy(T, X)
        assert abs(orc.temperature_con(e, X) - T) < 1e-6 * T
