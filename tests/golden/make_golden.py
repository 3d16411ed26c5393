"""Write the golden fixtures under tests/golden/ from the reference's own Fortran code.

Runs only in the build container, where the reference tree exists: it drives
oracle/_ref/ref_harness (built by ``make -C oracle ref`` from the reference sources in place) on
seeded inputs and stores inputs + outputs as small .npz files.  Nothing here travels to or runs on
the GPU box; the fixtures do.

    python tests/golden/make_golden.py            # all fixtures
    python tests/golden/make_golden.py recon run  # a subset
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")
CHEM_DIR = "/root/reference/WORK/KineticModels/Chem_bin/GRI3.0_CH4"
SEED = 20261202


def harness(cmd: str, ints=(), dbls=()):
    """Run one harness sub-command; returns (ints, doubles) of its output file."""
    ints = np.asarray(ints, dtype=np.int32).ravel()
    dbls = np.asarray(dbls, dtype=np.float64).ravel()
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([ints.size, dbls.size], np.int32).tobytes())
            f.write(ints.tobytes())
            f.write(dbls.tobytes())
        subprocess.run([HARNESS, cmd, CHEM_DIR, fin, fout], check=True,
                       stdout=subprocess.DEVNULL)
        raw = open(fout, "rb").read()
        names = None
        if os.path.exists(fout + ".names"):
            names = [l.strip() for l in open(fout + ".names")]
    ni, nd = np.frombuffer(raw[:8], np.int32)
    iv = np.frombuffer(raw[8:8 + 4 * ni], np.int32).copy()
    dv = np.frombuffer(raw[8 + 4 * ni:8 + 4 * ni + 8 * nd], np.float64).copy()
    return iv, dv, names


def save(name: str, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)")


# ----------------------------------------------------------------------------------------------
_MECH = {}


def mech():
    if not _MECH:
        iv, dv, names = harness("mech", dbls=[0.0])
        _MECH.update(iv=iv, dv=dv, names=names)
    return _MECH


def species():
    return mech()["names"]


def weights():
    m = mech()
    kk = len(m["names"])
    return m["dv"][-2 * kk:-kk]


def gen_mech():
    m = mech()
    save("mech", iv=m["iv"], dv=m["dv"], names=np.array(m["names"]))


def random_mixtures(rng, n, kk, zero_frac=0.4, neg=False):
    """Dirichlet-ish compositions with many exact zeros; optional tiny negatives."""
    y = rng.gamma(0.3, 1.0, size=(n, kk))
    y[rng.random((n, kk)) < zero_frac] = 0.0
    y[:, 0] += 1e-3
    y /= y.sum(axis=1, keepdims=True)
    if neg:
        mask = rng.random((n, kk)) < 0.05
        y[mask] = -rng.random(mask.sum()) * 1e-6
    return y


def air_y(kk):
    names = species()
    y = np.zeros(kk)
    y[names.index("O2")] = 0.233
    y[names.index("N2")] = 0.767
    return y


def ch4_air_x(kk):
    names = species()
    x = np.zeros(kk)
    x[names.index("CH4")] = 0.0950570342
    x[names.index("O2")] = 0.1901140684
    x[names.index("N2")] = 0.7148288973
    return x


def gen_recon():
    rng = np.random.default_rng(SEED + 1)
    rows = [rng.standard_normal((256, 8)),
            np.cumsum(rng.standard_normal((64, 8)) * 1e-4, axis=1) + 1.0,  # near-smooth
            np.repeat(rng.standard_normal((16, 1)), 8, axis=1),             # constant
            np.outer(rng.standard_normal(16), np.arange(8.0)) + rng.standard_normal((16, 1)),
            np.where(np.arange(8) < rng.integers(1, 8, (64, 1)), 1.0, 0.125),  # jumps
            rng.uniform(250.0, 3000.0, (32, 8)),                            # temperatures
            rng.uniform(0.0, 1e-7, (32, 8)),                                # eps-scale
            np.zeros((4, 8))]
    s = np.vstack(rows)
    _, out, _ = harness("recon", [s.shape[0]], s)
    save("recon", s=s, out=out.reshape(s.shape[0], 18))


def thermo_rows(rng, kk):
    t = np.concatenate([rng.uniform(200.0, 3500.0, 40),
                        [1000.0, np.nextafter(1000.0, 0), np.nextafter(1000.0, 2000.0),
                         300.0, 298.15, 2000.0, 2600.0, 1000.0]])
    n = t.size
    y = random_mixtures(rng, n, kk, neg=True)
    y[3] = air_y(kk)
    y[7] = air_y(kk)
    rho = np.exp(rng.uniform(np.log(0.05), np.log(30.0), n))
    return t, rho, y


def gen_thermo():
    rng = np.random.default_rng(SEED + 2)
    kk = len(species())
    t, rho, y = thermo_rows(rng, kk)
    inp = np.column_stack([t, rho, y])
    _, out, _ = harness("thermo", [t.size], inp)
    save("thermo", T=t, rho=rho, Y=y, out=out.reshape(t.size, 6 + 7 * kk))


def gen_eos():
    rng = np.random.default_rng(SEED + 3)
    kk = len(species())
    n = 48
    nequ = 3 + kk
    y = random_mixtures(rng, n, kk, neg=True) * rng.uniform(0.98, 1.02, (n, 1))
    qf = np.column_stack([np.exp(rng.uniform(np.log(0.05), np.log(30.0), n)),
                          rng.uniform(-800.0, 800.0, n), rng.uniform(250.0, 3200.0, n), y])
    # conserved rows: build from the primitive rows with a consistent energy, then perturb
    _, out0, _ = harness("eos", [n], np.concatenate([qf.ravel(), np.ones(n * nequ)]))
    qc = out0[:n * nequ].reshape(n, nequ).copy()
    qc[:, 2] *= rng.uniform(0.9, 1.1, n)
    qc[:, 3:] *= rng.uniform(0.995, 1.005, (n, kk))
    _, out, _ = harness("eos", [n], np.concatenate([qf.ravel(), qc.ravel()]))
    save("eos", qf=qf, Qc=qc, qf2Qc=out[:n * nequ].reshape(n, nequ),
         Qc2qf=out[n * nequ:].reshape(n, nequ))


def gen_rates():
    rng = np.random.default_rng(SEED + 4)
    kk = len(species())
    iv = mech()["iv"]
    ii = int(iv[2])
    n = 40
    t = np.concatenate([rng.uniform(300.0, 3000.0, n - 4), [1000.0, 1300.0, 2200.0, 700.0]])
    y = random_mixtures(rng, n, kk)
    y[-1] = 0.0
    y[-1, species().index("CH4")] = 0.0551
    y[-1, species().index("O2")] = 0.2200
    y[-1, species().index("N2")] = 0.7249
    p_atm = np.exp(rng.uniform(np.log(0.1), np.log(50.0), n))
    w = weights()
    wbar = 1.0 / (y / w).sum(axis=1)
    rho = p_atm * 101325.0 * wbar * 1e-3 / (8.31451 * t)  # kg/m^3
    inp = np.column_stack([rho, t, y])
    _, out, _ = harness("rates", [n], inp)
    save("rates", rho=rho, T=t, Y=y, out=out.reshape(n, 3 * kk + 4 * ii + 1))
    rows = []
    for r in range(16):
        y0 = random_mixtures(rng, 1, kk)[0]
        y1 = y0 + rng.standard_normal(kk) * 1e-4 * (y0 > 0)
        rows.append(np.concatenate([[rho[r], 10.0 ** rng.uniform(-8, -6)], y0, y1]))
    rows = np.array(rows)
    _, hout, _ = harness("hrr", [16], rows)
    save("hrr", rows=rows, hrr=hout)


def gen_trans():
    rng = np.random.default_rng(SEED + 5)
    kk = len(species())
    n = 32
    t = np.concatenate([rng.uniform(300.0, 3000.0, n - 2), [1000.0, 300.0]])
    p = np.exp(rng.uniform(np.log(1e4), np.log(5e6), n))
    x = random_mixtures(rng, n, kk)
    x[-1] = ch4_air_x(kk)
    _, out, _ = harness("trans", [n], np.column_stack([t, p, x]))
    save("trans", T=t, p=p, X=x, out=out.reshape(n, 2 + kk))


def flux_state(rng, n, lv, kk):
    """Primitive rows (rho, u, T, Y...) for cells 1-lv .. n+lv: smooth waves, a contact, a shock,
    supersonic patches of both signs."""
    m = n + 2 * lv
    x = np.linspace(0.0, 1.0, m)
    rho = 1.0 + 0.3 * np.sin(6 * x) + np.where(x > 0.55, 1.5, 0.0)
    u = 250.0 * np.sin(9 * x) + np.where(x < 0.2, 700.0, 0.0) - np.where(x > 0.85, 900.0, 0.0)
    t = 900.0 + 600.0 * np.cos(5 * x) + np.where(x > 0.4, 400.0, 0.0)
    y = random_mixtures(rng, m, kk) * (1.0 + 1e-3 * rng.standard_normal((m, 1)))
    y[: m // 3] = 0.7 * y[: m // 3] + 0.3 * y[0]
    return np.column_stack([rho, u, t, y])


FLUX_CASES = [  # (flag_acu, flag_weno, flag_muscl, select_riemann)
    (2, 2, 0, 2), (2, 3, 0, 2), (2, 1, 0, 2), (2, 4, 0, 2),
    (1, 0, 2, 2), (1, 0, 1, 2), (1, 0, 3, 2), (1, 0, 4, 2),
    (0, 0, 0, 2), (2, 2, 0, 0), (2, 2, 0, 1), (1, 0, 2, 1),
]


def gen_flux():
    rng = np.random.default_rng(SEED + 6)
    kk = len(species())
    n, lv, dx = 40, 4, 1e-3
    qf = flux_state(rng, n, lv, kk)
    outs = []
    for case in FLUX_CASES:
        _, out, _ = harness("flux", [n, lv, *case], np.concatenate([[dx], qf.ravel()]))
        outs.append(out)
    nequ = 3 + kk
    outs = np.array(outs)
    save("flux", qf=qf, n=n, lv=lv, dx=dx, cases=np.array(FLUX_CASES),
         Fplus=outs[:, : nequ * (n + 1)].reshape(len(FLUX_CASES), n + 1, nequ),
         godunov=outs[:, nequ * (n + 1):].reshape(len(FLUX_CASES), n + 1, 3))


def sod_cells(kk, ncell, dx):
    xm = (np.arange(1, ncell + 1) - 0.5) * dx
    names = species()
    w = weights()
    y = air_y(kk)
    x = (y / w) / (y / w).sum()
    rows = np.zeros((ncell, 3 + kk))
    rows[:, 0] = 300.0
    rows[:, 1] = np.where(xm <= 0.5, 1.0e6, 1.0e5)
    rows[:, 3:] = x
    del names
    return rows


def hot_ch4_cells(kk, ncell, dx, nhot):
    x = ch4_air_x(kk)
    rows = np.zeros((ncell, 3 + kk))
    rows[:, 0] = 300.0
    rows[:nhot, 0] = 2000.0
    rows[:, 1] = 101325.0
    rows[:, 3:] = x
    return rows


RUNS = {  # name: (nblk, ncell, nsteps, lvir, acu, weno, muscl, riemann, time, cfl_flag, dx, cfl, dt)
    "sod_weno5js_hllc": (1, 1000, 200, 3, 2, 2, 0, 2, 1, 1, 1e-3, 0.5, 0.0),
    "sod_weno5z_4blk": (4, 250, 20, 3, 2, 3, 0, 2, 1, 1, 1e-3, 0.5, 0.0),
    "sod_muscl_100blk": (100, 10, 40, 2, 1, 0, 2, 2, 1, 1, 1e-3, 0.4, 0.0),
    "sod_euler_hll": (2, 50, 10, 3, 2, 2, 0, 1, 0, 1, 2e-2, 0.3, 0.0),
    "sod_rk4_lxf": (2, 50, 10, 2, 1, 0, 2, 0, 2, 0, 2e-2, 0.0, 2e-6),
    "ch4_hot_weno5": (2, 64, 10, 3, 2, 2, 0, 2, 1, 1, 25e-6, 0.5, 0.0),
}


def gen_run(which=None):
    kk = len(species())
    for name, (nblk, ncell, nsteps, lv, acu, wen, mus, riem, tim, cflf, dx, cfl, dt) in RUNS.items():
        if which and name not in which:
            continue
        ntot = nblk * ncell
        if name.startswith("sod"):
            cells = sod_cells(kk, ntot, dx) if ntot * dx == 1.0 else sod_cells_scaled(kk, ntot, dx)
        else:
            cells = hot_ch4_cells(kk, ntot, dx, 16)
        _, out, _ = harness("run", [nblk, ncell, nsteps, lv, acu, wen, mus, riem, tim, cflf],
                            np.concatenate([[dx, cfl, dt], cells.ravel()]))
        nequ = 3 + kk
        per = nblk * (ncell + 2 * lv) * nequ
        secs, dts = out[0], out[1:1 + nsteps]
        st = out[1 + nsteps:].reshape(4, nblk, ncell + 2 * lv, nequ)
        save("run_" + name, cells=cells, params=np.array([nblk, ncell, nsteps, lv, acu, wen, mus,
                                                          riem, tim, cflf]),
             dx=dx, cfl=cfl, dt=dt, dts=dts, seconds=secs,
             Qc0=st[0], qf0=st[1], Qc=st[2], qf=st[3])
        del per


def sod_cells_scaled(kk, ntot, dx):
    rows = sod_cells(kk, ntot, dx)
    xm = (np.arange(1, ntot + 1) - 0.5) * dx
    rows[:, 1] = np.where(xm <= 0.5 * ntot * dx, 1.0e6, 1.0e5)
    return rows


def main(argv):
    want = set(argv[1:])
    jobs = {"mech": gen_mech, "recon": gen_recon, "thermo": gen_thermo, "eos": gen_eos,
            "rates": gen_rates, "trans": gen_trans, "flux": gen_flux}
    for k, f in jobs.items():
        if not want or k in want:
            f()
    runs = [w for w in want if w in RUNS]
    if not want or "run" in want or runs:
        gen_run(runs or None)


if __name__ == "__main__":
    main(sys.argv)
