"""The Fortran binding (fortrancfd_amd/fortran/fcfd.f90): a Fortran program drives the reference's
loop (calc_dt, solve_fluid, solve_source, refresh) through ISO_C_BINDING calls into libfcfd_hip.so.
Inert runs must reproduce the reference's golden runs bit for bit, like the Python binding does; a
reacting run (SRC/main.f90:142-145, Lie split) is checked against the oracle to 1e-12."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import REPO, golden

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(REPO, "tests", "fortran", "_build", "sod_driver")


@pytest.mark.parametrize("name", ["sod_weno5js_hllc", "sod_rk4_lxf"])
def test_fortran_driver_reproduces_reference_run(tmp_path, name):
    from fortrancfd_amd.lib import GRI30
    assert os.path.exists(DRIVER), "build the Fortran test driver first (__graft_entry__.build())"
    g = golden("run_" + name)
    nblk, ncell, ns, lv, acu, wen, mus, riem, tim, cflf = [int(x) for x in g["params"]]
    ints = np.array([nblk, ncell, ns, tim, riem, acu, mus, wen, cflf, 0], np.int32)
    dbl = np.array([float(g["dx"]), float(g["cfl"]), float(g["dt"]) if "dt" in g else 0.0])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(ints.tobytes())
        f.write(dbl.tobytes())
        f.write(np.ascontiguousarray(g["cells"], np.float64).tobytes())
    print(f"fortran driver {name}", flush=True)
    r = subprocess.run([DRIVER, GRI30, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, np.float64)
    dts = raw[:ns]
    rest = raw[ns:].reshape(nblk, 2, ncell + 2 * lv, -1)
    np.testing.assert_array_equal(dts, g["dts"])
    np.testing.assert_array_equal(rest[:, 0], g["Qc"])
    np.testing.assert_array_equal(rest[:, 1], g["qf"])


def test_fortran_driver_reacting_run_vs_oracle(tmp_path):
    """Three reacting steps (hot spot in stoichiometric CH4/air) through the Fortran binding,
    fcfd_solve_source included, against the oracle's Lie-split steps: dt, Qc, qf, HRR, HRR_dot and
    the integrator counters of the last call."""
    from fortrancfd_amd.lib import GRI30
    from oracle.pyoracle import Domain, Oracle
    from tests.test_gpu_chem import REL, Y_FLOOR, hot_spot_cells, hrr_errors
    orc = Oracle()
    cells, nblk, ncell, dx = hot_spot_cells(orc.kk)
    ns, cfl = 3, 0.5
    ints = np.array([nblk, ncell, ns, 1, 2, 2, 0, 2, 1, 1], np.int32)
    dbl = np.array([dx, cfl, 0.0])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(ints.tobytes())
        f.write(dbl.tobytes())
        f.write(np.ascontiguousarray(cells, np.float64).tobytes())
    print("fortran driver, reacting", flush=True)
    r = subprocess.run([DRIVER, GRI30, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    d = Domain(orc, cells, nblk, ncell, dx)
    dts = []
    for _ in range(ns):
        dt = d.calc_dt(cfl)
        dts.append(dt)
        d.solve_fluid(dt)
        hrr, hrd, nf = d.solve_source(dt)
        assert nf == 0
        d.refresh()
    lv, nequ, n = d.lv, orc.nequ, nblk * ncell
    raw = open(fout, "rb").read()
    o = 0
    g_dts = np.frombuffer(raw, np.float64, ns, o); o += 8 * ns
    m = nblk * 2 * (ncell + 2 * lv) * nequ
    blocks = np.frombuffer(raw, np.float64, m, o).reshape(nblk, 2, ncell + 2 * lv, nequ); o += 8 * m
    g_hrr = np.frombuffer(raw, np.float64, n, o); o += 8 * n
    g_hrd = np.frombuffer(raw, np.float64, n, o); o += 8 * n
    g_st = np.frombuffer(raw, np.int32, 4 * n, o).reshape(n, 4)
    np.testing.assert_allclose(g_dts, dts, rtol=1e-14, atol=0)
    assert (g_st[:, :3] == d.chem_stats()[:, :3]).all()
    # conserved fields: rho and rho E relative to their largest magnitude over the domain; rho u
    # (near rest, a difference of nearly equal pressure fluxes) relative to the acoustic momentum
    # scale max(rho a); partial densities relative to max(rho Y_k, rho Y_FLOOR)
    Qg, Qo = blocks[:, 0], d.Qc
    for n in (0, 2):
        assert np.abs(Qg[..., n] - Qo[..., n]).max() <= REL * np.abs(Qo[..., n]).max(), n
    rho, T, Y = d.qf[..., 0], d.qf[..., 2], d.qf[..., 3:]
    a = np.sqrt(1.4 * 8314.51 * T * (Y / orc.weights()).sum(axis=-1))      # sound speed scale, m/s
    assert np.abs(Qg[..., 1] - Qo[..., 1]).max() <= REL * (rho * a).max()
    floor = Y_FLOOR * Qo[..., :1]
    assert (np.abs(Qg[..., 3:] - Qo[..., 3:]) / np.maximum(np.abs(Qo[..., 3:]), floor)).max() <= REL
    np.testing.assert_allclose(blocks[:, 1, :, 2], d.qf[..., 2], rtol=REL, atol=0)
    Yg, Yo = blocks[:, 1, :, 3:], d.qf[..., 3:]
    assert (np.abs(Yg - Yo) / np.maximum(np.abs(Yo), Y_FLOOR)).max() <= REL
    eh, ed = hrr_errors(orc, d, dts[-1], g_hrr, g_hrd, hrr, hrd, field=True)
    assert eh <= REL and ed <= REL


def test_fortran_flame_driver_vs_oracle(tmp_path):
    """The spherical-flame configuration from Fortran: a stretched mesh (two refinement levels on the
    right), spherical geometry, outflow, radiation of the deck's rad_params_4sp.txt and its ignition
    source, chemistry, four driver steps with the time advanced before the fluid step; the output
    rows (fcfd_eos_fields) against the oracle's, within 1e-12 of each row's largest magnitude (the
    velocity of the sound speed, the kinetic energy of the total energy)."""
    from fortrancfd_amd import inputs
    from fortrancfd_amd.lib import GRI30
    from oracle.pyoracle import Domain, Oracle
    from tests.test_gpu_aux import deck
    from tests.test_gpu_chem import REL
    from tests.test_gpu_configs import ch4_air_cells
    orc = Oracle()
    nm = list(golden("mech")["names"])
    kk = len(nm)
    nfine, ncell, smr, dx, cfl, ns = 2, 16, (1, 0, 3, 1), 2e-5, 0.4, 4
    nblk = nfine + 2
    cells = ch4_air_cells(kk, nblk * ncell, [(0, 12, 2100.0)])
    base = ("CH4_GRI3.0_SphericalFlame_radiation",)
    idx, pm = inputs.read_radiation_params(deck(*base, "rad_params_4sp.txt"), nm)
    ig = inputs.read_igsource(deck(*base, "igsource.inp"))
    ints = np.array([nfine, ncell, *smr, ns, 1, len(idx)], np.int32)
    dbl = np.array([dx, cfl, 1.0, ig["E_mJ"][0], ig["r_um"][0], ig["t_us"][0], ig["t0_us"][0]])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(ints.tobytes())
        f.write(dbl.tobytes())
        f.write(np.ascontiguousarray(idx, np.int32).tobytes())
        f.write(np.ascontiguousarray(pm, np.float64).tobytes())
        f.write(np.ascontiguousarray(cells, np.float64).tobytes())
    print("fortran flame driver", flush=True)
    drv = os.path.join(REPO, "tests", "fortran", "_build", "flame_driver")
    r = subprocess.run([drv, GRI30, str(fin), str(fout)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(fout, "rb").read()
    nb = int(np.frombuffer(raw[:4], np.int32)[0])
    assert nb == nblk
    vals = np.frombuffer(raw[4:], np.float64)
    dts, fields = vals[:ns], vals[ns:].reshape(15 + kk, nblk * ncell)
    d = Domain(orc, cells, nfine, ncell, dx, smr=smr, is_spherical=1)
    d.set_flow_boundary(0, 1, p_out_atm=1.0)
    d.set_radiation(idx, pm)
    d.set_ignition(1, **ig)
    dts_o = [d.step(cfl, react=True) for _ in range(ns)]
    np.testing.assert_allclose(dts, dts_o, rtol=1e-13)
    f_o = d.eos_fields()
    # the velocity near rest (a difference of nearly equal pressure fluxes) on the acoustic scale,
    # the kinetic energy on the energy scale; every other row on its own largest magnitude
    a = np.sqrt(1.4 * f_o[3] / f_o[0])                  # sqrt(gamma p / rho)
    scales = {1: a.max(), 5: np.abs(f_o[4]).max()}
    for r_ in range(15 + kk):
        if r_ in (7, 8):                                 # HRR rows: not requested by the driver
            continue
        sc = scales.get(r_, max(np.abs(f_o[r_]).max(), 1e-300 if r_ < 15 else 1e-12))
        assert np.abs(fields[r_] - f_o[r_]).max() <= REL * sc, r_
