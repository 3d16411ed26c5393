"""HIP fluid path (libfcfd_hip.so through the C ABI) against the reference's golden vectors and the
CPU oracle.  The fluid kernels are compiled without FMA contraction and follow the reference's
operation order, so the expected agreement is bit-for-bit; the tolerance written in each test is
zero unless stated."""
import numpy as np
import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mech():
    from fortrancfd_amd.lib import Mechanism
    return Mechanism()


def test_recon_golden():
    from fortrancfd_amd.lib import recon
    g = golden("recon")
    np.testing.assert_array_equal(recon(g["s"]), g["out"])


def test_eos_golden(mech):
    g = golden("eos")
    np.testing.assert_array_equal(mech.qfToQc(g["qf"]), g["qf2Qc"])
    np.testing.assert_array_equal(mech.QcToqf(g["Qc"]), g["Qc2qf"])


@pytest.mark.parametrize("ci", range(12))
def test_flux_golden(mech, ci):
    g = golden("flux")
    acu, wen, mus, riem = (int(x) for x in g["cases"][ci])
    print(f"flux case {ci}: acu={acu} weno={wen} muscl={mus} riemann={riem}", flush=True)
    F, G = mech.calc_flux_advection(g["qf"], int(g["n"]), int(g["lv"]), float(g["dx"]),
                                    flag_acu=acu, flag_weno=wen, flag_muscl=mus,
                                    select_riemann=riem)
    np.testing.assert_array_equal(F, g["Fplus"][ci])
    np.testing.assert_array_equal(G, g["godunov"][ci])


def _solver_for(mech, g):
    from fortrancfd_amd.lib import Solver
    nblk, ncell, ns, lv, acu, wen, mus, riem, tim, cflf = [int(x) for x in g["params"]]
    s = Solver(mech, nblk, ncell, float(g["dx"]), flag_acu=acu, flag_weno=wen, flag_muscl=mus,
               select_riemann=riem, select_time=tim)
    assert s.lv == lv
    return s, ns, cflf


GPU_RUNS = ["sod_weno5js_hllc", "sod_weno5z_4blk", "sod_muscl_100blk", "sod_euler_hll",
            "sod_rk4_lxf", "ch4_hot_weno5"]


@pytest.mark.parametrize("name", GPU_RUNS)
def test_runs_golden(mech, name):
    """Whole driver loop (calc_dt, solve_fluid, refresh) from the reference's initial state, every
    step's dt and the final Qc / qf of every block (ghosts included) bit-for-bit."""
    g = golden("run_" + name)
    print(f"run {name}", flush=True)
    s, ns, cflf = _solver_for(mech, g)
    s.init_cells(g["cells"])
    Qc0, qf0 = s.download()
    np.testing.assert_array_equal(Qc0, g["Qc0"])
    np.testing.assert_array_equal(qf0, g["qf0"])
    dts = []
    for _ in range(ns):
        dt = s.calc_dt(float(g["cfl"])) if cflf else float(g["dt"])
        dts.append(dt)
        s.solve_fluid(dt)
        s.refresh()
    np.testing.assert_array_equal(np.array(dts), g["dts"])
    Qc, qf = s.download()
    np.testing.assert_array_equal(Qc, g["Qc"])
    np.testing.assert_array_equal(qf, g["qf"])


@pytest.mark.parametrize("name", ["sod_weno5z_4blk", "sod_rk4_lxf", "sod_euler_hll"])
def test_advance_matches_stepwise(mech, name):
    """Device-resident stepping (dt on device, captured graph) equals the host-driven loop with
    the same CFL steps, bit for bit; for the CFL-stepped reference runs both equal the reference."""
    g = golden("run_" + name)
    a, ns, cflf = _solver_for(mech, g)
    b, _, _ = _solver_for(mech, g)
    a.init_cells(g["cells"])
    b.init_cells(g["cells"])
    cfl = float(g["cfl"]) if cflf else 0.5          # fixed-dt reference runs: step at CFL 0.5
    t = a.advance(cfl, ns, 0.0)
    tb = 0.0
    for _ in range(ns):
        dt = b.calc_dt(cfl)
        tb += dt
        b.solve_fluid(dt)
    assert t == pytest.approx(tb, rel=1e-14)
    np.testing.assert_array_equal(a.download()[0], b.download()[0])
    if cflf:
        np.testing.assert_array_equal(a.download()[0], g["Qc"])


def test_upload_roundtrip_and_oracle(mech, orc):
    """Upload an oracle state, step both three times, compare (seeded reacting composition)."""
    from oracle.pyoracle import Domain
    from fortrancfd_amd.lib import Solver
    rng = np.random.default_rng(7)
    nblk, ncell, dx = 3, 48, 25e-6
    kk = mech.kk
    cells = np.zeros((nblk * ncell, 3 + kk))
    cells[:, 0] = rng.uniform(300.0, 2200.0, nblk * ncell)
    cells[:, 1] = 101325.0 * rng.uniform(0.8, 1.2, nblk * ncell)
    cells[:, 2] = rng.uniform(-30.0, 30.0, nblk * ncell)
    cells[:, 3:] = rng.dirichlet(np.full(kk, 0.3), nblk * ncell)
    d = Domain(orc, cells, nblk, ncell, dx, flag_acu=2, flag_weno=3, select_riemann=2)
    s = Solver(mech, nblk, ncell, dx, flag_acu=2, flag_weno=3, select_riemann=2)
    s.upload(d.Qc, d.qf)
    Qc, qf = s.download()
    np.testing.assert_array_equal(Qc, d.Qc)
    for _ in range(3):
        dt_o, dt_g = d.calc_dt(0.4), s.calc_dt(0.4)
        assert dt_o == dt_g
        d.solve_fluid(dt_o)
        d.refresh()
        s.solve_fluid(dt_g)
    Qc, qf = s.download()
    np.testing.assert_array_equal(Qc, d.Qc)
    np.testing.assert_array_equal(qf, d.qf)


def test_large_domain_conservation(mech):
    """65536 cells in 16 blocks (the coupled-run shape), stoichiometric CH4/air with three hot
    spots, 20 steps: reflective walls conserve mass, every species' mass and total energy to
    round-off; no NaN; temperatures stay in the Newton clamp."""
    from fortrancfd_amd.lib import Solver
    kk = mech.kk
    nblk, ncell, dx = 16, 4096, 25e-6
    n = nblk * ncell
    rng = np.random.default_rng(20261202)
    names = list(golden("mech")["names"])
    x = np.zeros(kk)
    x[names.index("CH4")] = 0.0950570342
    x[names.index("O2")] = 0.1901140684
    x[names.index("N2")] = 0.7148288973
    cells = np.zeros((n, 3 + kk))
    cells[:, 0] = 300.0
    cells[:, 1] = 101325.0
    cells[:, 3:] = x
    for c in rng.integers(0, n - 32, 3):
        cells[c:c + 32, 0] = rng.uniform(1800.0, 2200.0)
    s = Solver(mech, nblk, ncell, dx, flag_acu=2, flag_weno=2, select_riemann=2)
    s.init_cells(cells)
    Qc0, _ = s.download()
    lv = s.lv
    t = s.advance(0.5, 20)
    Qc, qf = s.download()
    assert t > 0 and np.isfinite(Qc).all()
    tot0 = Qc0[:, lv:-lv, :].sum(axis=(0, 1))
    tot1 = Qc[:, lv:-lv, :].sum(axis=(0, 1))
    for v in [0, 2] + list(range(3, 3 + kk)):
        assert abs(tot1[v] - tot0[v]) <= 1e-11 * max(abs(tot0[v]), 1e-300), v
    T = qf[:, lv:-lv, 2]
    assert T.min() > 250.0 and T.max() < 2300.0
