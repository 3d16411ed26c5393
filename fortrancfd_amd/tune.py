"""Launch-shape tuner: times the candidate shapes of the hot kernels on this GPU and writes the
table that the library reads at start-up (fcfd_set_launch / fcfd_load_tuning).

    python -m fortrancfd_amd.tune [--out fortrancfd_amd/data/tuned_launch.txt] [--steps N]

Candidates and the workloads they are timed on (GRI-3.0):
  * flux_variant (WENO5 advection flux: 64 threads at two or one waves per SIMD, 128 threads) and
    stage_variant (stage update: 64 / two, 64 / one, 128 / one, 256 / one) on inert hydrodynamics,
    64 blocks x 4096 cells, WENO5-JS + HLLC + SSP-RK3, kernel times from the context's HIP events;
  * chem_wg_per_cu (implicit-integrator workgroups per CU) on a long stiff list: 2^16 independent
    cells of the chemistry batch generator (fortrancfd_amd.workloads.chemistry_batch);
  * chem_four_wave_factor (the list length below which the integrator gives each cell four waves)
    and chem_lu (the four-wave mode's LU: step by step through LDS, in 16-column panels with the
    FP64 matrix cores, or with the matrix in registers; the factors are bitwise the same) on whole
    steps of the coupled 65536-cell domain: the run continued for 200 steps from its state after
    --steps Strang steps, per candidate (chem_lu first, at the default factor).
A candidate replaces the default only when it is more than 2 % faster.
Each candidate is timed three times and scored by its median; the table records the winners and,
as comments, every candidate's time.
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import time

import numpy as np

from .lib import GRI30, Mechanism, Solver, _check, lib

PKG = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(PKG, "data", "tuned_launch.txt")


def set_launch(key, nsp, value):
    _check(lib().fcfd_set_launch(key.encode(), C.c_int(nsp), C.c_int(value)))


def median(xs):
    return float(np.median(np.asarray(xs)))


def pick(res, default, margin=0.02):
    """The fastest candidate, or the default unless another beats it by more than margin."""
    best = min(res, key=res.get)
    return best if res[best] < (1.0 - margin) * res[default] else default


def tune_fluid(mech, log):
    from .workloads import coupled_cells
    nblk, ncell = 64, 4096
    cells = coupled_cells(mech, nblk, ncell, nhot=0)
    out = {}
    for key, kid, cands in (("flux_variant", 0, (0, 1, 2)), ("stage_variant", 1, (0, 1, 2, 3))):
        res = {}
        for v in cands:
            set_launch(key, mech.kk, v)
            s = Solver(mech, nblk, ncell, 25e-6, flag_acu=2, flag_weno=2, select_riemann=2, select_time=1)
            s.init_cells(cells)
            s.set_timing(True)
            s.advance(0.5, 2)                                   # warm-up
            reps = []
            for _ in range(3):
                ms0, n0 = s.timing(kid)
                s.advance(0.5, 2)
                ms1, n1 = s.timing(kid)
                reps.append((ms1 - ms0) / max(n1 - n0, 1))
            res[v] = median(reps)
            log(f"{key}={v}: {res[v]:.4f} ms per launch")
            del s
        best = pick(res, 0)
        set_launch(key, mech.kk, best)
        out[key] = (best, res)
    return out


def time_batch(mech, n, seed, key, cands, log, what):
    import torch
    from .workloads import chemistry_batch
    rho, dt, T, Y = chemistry_batch(mech, n, seed=seed)
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    res = {}
    for v in cands:
        set_launch(key, mech.kk, v)
        reps = []
        for _ in range(3):
            tr, tdt, Tf, Yf = d(rho), d(dt), d(T), d(Y.T)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            mech.chem_batch_device(tr, tdt, Tf, Yf)
            torch.cuda.synchronize()
            reps.append((time.perf_counter() - t0) * 1e3)
        res[v] = median(reps)
        log(f"{key}={v}: {res[v]:.2f} ms for {n} {what}")
    return res


def tune_chem_long(mech, log, n=1 << 16):
    from .workloads import SEED
    res = time_batch(mech, n, SEED + 11, "chem_wg_per_cu", (3, 4, 5), log, "batch cells")
    best = pick(res, 5)
    set_launch("chem_wg_per_cu", mech.kk, best)
    return best, res


def tune_chem_short(mech, log, steps, window=200):
    """chem_four_wave_factor and chem_lu on whole coupled steps: the coupled 65536-cell domain
    (Strang, HRR every step) run to `steps`, then `window` steps from that saved state per candidate
    (ms per step, median of three).  The window takes in the stiff lists of every length the run
    produces -- the short lists of the hot spots and the periodic bursts of a thousand or more cold
    cells -- so the factor is chosen for the mixture the production step sees."""
    from .workloads import coupled_cells
    cells = coupled_cells(mech)
    s = Solver(mech, 16, 4096, 25e-6, flag_acu=2, flag_weno=2, select_riemann=2, select_time=1,
               do_reaction=1)
    s.set_step(strang=True, hrr_each_step=True)
    s.init_cells(cells)
    t = s.advance(0.5, steps)
    Qc, qf = s.download()

    def per_step():
        reps = []
        for _ in range(3):
            s.upload(Qc, qf)                 # the same start, step-size memory cleared
            t0 = time.perf_counter()
            s.advance(0.5, window, t)
            reps.append((time.perf_counter() - t0) * 1e3 / window)
        return median(reps)

    # the four-wave mode's factorisation first (step by step through LDS, blocked on the matrix
    # cores, or in registers: the same factors), at the default factor
    lu = {}
    for v in (0, 1, 2):
        set_launch("chem_lu", mech.kk, v)
        lu[v] = per_step()
        log(f"chem_lu={v}: {lu[v]:.3f} ms per coupled step (four-wave factor 8)")
    best_lu = pick(lu, 2)
    set_launch("chem_lu", mech.kk, best_lu)
    res = {}
    for v in (0, 1, 2, 4, 8):
        set_launch("chem_four_wave_factor", mech.kk, v)
        res[v] = per_step()
        log(f"chem_four_wave_factor={v}: {res[v]:.3f} ms per coupled step (steps {steps}-{steps + window}, chem_lu {best_lu})")
    best = pick(res, 8)
    set_launch("chem_four_wave_factor", mech.kk, best)
    return best, res, best_lu


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--steps", type=int, default=2000, help="coupled steps before the timed window")
    args = ap.parse_args(argv)
    import torch
    assert torch.cuda.is_available(), "the tuner needs a GPU"
    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    mech = Mechanism(GRI30)
    lines = []

    def log(msg):
        print(msg, flush=True)
        lines.append("# " + msg)

    # start from the defaults, whatever table is installed
    for key, v in (("flux_variant", 0), ("stage_variant", 0), ("chem_wg_per_cu", 5), ("chem_four_wave_factor", 8),
                   ("chem_lu", 2)):
        set_launch(key, mech.kk, v)
    fl = tune_fluid(mech, log)
    wg, _ = tune_chem_long(mech, log)
    fw, _, lu = tune_chem_short(mech, log, args.steps)
    table = [f"{arch} {mech.kk} flux_variant {fl['flux_variant'][0]}",
             f"{arch} {mech.kk} stage_variant {fl['stage_variant'][0]}",
             f"{arch} {mech.kk} chem_wg_per_cu {wg}",
             f"{arch} {mech.kk} chem_four_wave_factor {fw}",
             f"{arch} {mech.kk} chem_lu {lu}"]
    with open(args.out, "w") as f:
        f.write("# Launch shapes chosen by python -m fortrancfd_amd.tune (median of three timings each).\n")
        f.write("# Lines: <arch> <species count> <key> <value>; read by libfcfd_hip.so at start-up.\n")
        for ln in lines:
            f.write(ln + "\n")
        for ln in table:
            f.write(ln + "\n")
    print("\n".join(table), flush=True)
    return table


if __name__ == "__main__":
    main()
