"""Benchmark of the per-timestep cell and interface update on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload coupled65536|inert65536|inert1m]

Prints one JSON line.  The default workload is the 65536-cell GRI-3.0 domain of the coupled run
(16 blocks x 4096 cells, dx = 25 um, stoichiometric CH4/air at 1 atm and 300 K with three seeded
32-cell hot spots at 1800-2200 K, WENO5-JS + HLLC + SSP-RK3, CFL 0.5, reflective walls), stepped
with its hydrodynamics on the GPU.  Synthetic data, FP64 throughout.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
SEED = 20261202


def ch4_air_cells(names, n, nhot, seed):
    kk = len(names)
    x = np.zeros(kk)
    x[names.index("CH4")] = 0.0950570342
    x[names.index("O2")] = 0.1901140684
    x[names.index("N2")] = 0.7148288973
    cells = np.zeros((n, 3 + kk))
    cells[:, 0] = 300.0
    cells[:, 1] = 101325.0
    cells[:, 3:] = x
    rng = np.random.default_rng(seed)
    for c in rng.integers(0, n - 32, nhot):
        cells[c:c + 32, 0] = rng.uniform(1800.0, 2200.0)
    return cells


def species_names(mech_dir):
    """Species symbols from chem.bin (16-character records of the species list)."""
    import struct
    raw = open(os.path.join(mech_dir, "chem.bin"), "rb").read()
    recs, o = [], 0
    while o < len(raw):
        n = struct.unpack("<i", raw[o:o + 4])[0]
        recs.append(raw[o + 4:o + 4 + n])
        o += 8 + n
    hdr = struct.unpack("<29i", recs[1][:116])
    mm, kk = hdr[3], hdr[4]
    per = 16 + 4 * mm + 4 + 4 + 8 + 4 + 3 * 8 + 2 * 7 * 8
    return [recs[3][k * per:k * per + 16].decode().strip() for k in range(kk)]


def reference_cpu_baseline(cells, nblk, ncell, dx, steps, cfl):
    """Times the reference Fortran path (oracle/_ref/ref_harness, amdflang -O2, one core) on a
    bounded sample of the same workload.  Returns None when the harness is not built."""
    harness = os.path.join(REPO, "oracle", "_ref", "ref_harness")
    if not os.path.exists(harness):
        return None
    from fortrancfd_amd.lib import GRI30
    ints = np.array([nblk, ncell, steps, 3, 2, 2, 0, 2, 1, 1], np.int32)
    dbls = np.concatenate([[dx, cfl, 0.0], cells.ravel()]).astype(np.float64)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([ints.size, dbls.size], np.int32).tobytes())
            f.write(ints.tobytes())
            f.write(dbls.tobytes())
        subprocess.run([harness, "run", GRI30, fin, fout], check=True, stdout=subprocess.DEVNULL)
        raw = open(fout, "rb").read()
    ni = int(np.frombuffer(raw[:4], np.int32)[0])
    secs = float(np.frombuffer(raw[8 + 4 * ni:16 + 4 * ni], np.float64)[0])
    return secs


WORKLOADS = {
    # name: (nblk, ncell, dx, hot spots)
    "inert65536": (16, 4096, 25e-6, 3),
    "inert1m": (256, 4096, 25e-6, 48),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="inert65536", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every block of the domain lives on one device and this "
              "build has no multi-GPU path (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    from fortrancfd_amd.lib import GRI30, Mechanism, Solver

    nblk, ncell, dx, nhot = WORKLOADS[args.workload]
    ntot = nblk * ncell
    names = species_names(GRI30)
    cells = ch4_air_cells(names, ntot, nhot, SEED)
    mech = Mechanism()
    s = Solver(mech, nblk, ncell, dx, flag_acu=2, flag_weno=2, select_riemann=2, select_time=1)
    s.init_cells(cells)
    cfl = 0.5
    t = s.advance(cfl, args.warmup)

    s.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t = s.advance(cfl, args.steps, t)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    flux_ms, flux_n = s.timing(0)
    stage_ms, stage_n = s.timing(1)
    dt_ms, dt_n = s.timing(2)
    s.set_timing(False)

    cell_steps = ntot * args.steps
    value = cell_steps / elapsed
    nequ = mech.nequ
    # flux kernel: reads the primitive state once (nequ doubles per cell) and writes one flux
    # vector per interface (nequ doubles); ncell+1 interfaces per block
    flux_bytes = 8.0 * nequ * (ntot + nblk * (ncell + 1))
    flux_avg_s = flux_ms / flux_n / 1e3
    achieved = flux_bytes / flux_avg_s / 1e9

    cpu = None
    if not args.no_cpu:
        csteps = args.cpu_steps
        secs = reference_cpu_baseline(cells, nblk, ncell, dx, csteps, cfl)
        if secs is not None:
            cpu = {"value": ntot * csteps / secs, "unit": "cell-steps/s", "cores": 1,
                   "kind": "reference",
                   "sample": f"{csteps} steps of the same {ntot}-cell domain on the reference "
                             f"Fortran (amdflang -O2, one core): {secs:.2f} s"}

    line = {
        "metric": "cell-steps/s (GRI-3.0 WENO5+HLLC+SSP-RK3 hydrodynamics)",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {nblk} blocks x {ncell} cells, GRI-3.0 "
                               f"CH4/air 1 atm 300 K, {nhot} hot spots, WENO5-JS+HLLC+SSP-RK3, "
                               f"CFL {cfl}, walls", "cells": ntot, "species": mech.kk},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_flux<53,WENO5JS>", "avg_launch_ms": flux_avg_s * 1e3,
                     "stage_end_avg_ms": stage_ms / max(stage_n, 1),
                     "cfl_avg_ms": dt_ms / max(dt_n, 1)},
        "cpu_baseline": cpu,
        "sim_time_s": t,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
