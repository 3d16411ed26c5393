"""Benchmark of the per-timestep cell and interface update on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload coupled65536|inert65536|inert1m|chem1m|surrogate1m]

Prints one JSON line.  The default workload is the coupled reacting run: 65536 cells in 16 blocks
of 4096 (dx = 25 um), stoichiometric CH4/air at 1 atm and 300 K with three seeded 32-cell hot spots
at 1800-2200 K, GRI-3.0 (53 species, 325 reactions), WENO5-JS + HLLC + SSP-RK3 at CFL 0.5 between
reflective walls, Strang-split chemistry (calc_dt, solve_source(dt/2), refresh, solve_fluid(dt),
solve_source(dt/2) with the HRR / HRR_dot diagnostics, refresh; replayed from a HIP graph).  K steps
are timed with the chemistry through the trained surrogate + FP64 fallback (`value`), then K steps
from the same initial state with the plain FP64 integrator; the line reports both rates, their ratio
and the final-state errors of the surrogate run against the FP64 run after all steps.  Synthetic
data, FP64 state throughout (BF16 / FP32 inside the network).  --lie times the reference driver's
own splitting (fluid, then a full chemistry step) instead.
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
FP64_PEAK_TF = 78.6
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
    # name: (nblk, ncell, dx, hot spots, chemistry)
    "coupled65536": (16, 4096, 25e-6, 3, True),
    "inert65536": (16, 4096, 25e-6, 3, False),
    "inert1m": (256, 4096, 25e-6, 48, False),
}
MAJOR = ("CH4", "O2", "N2", "CO2", "H2O", "CO", "H2")
TIMERS = {0: "k_flux", 1: "k_stage_end", 2: "k_wave_speed+k_dt_finalize", 3: "k_chem (SDIRK)",
          4: "k_qf_to_qc_interior", 5: "surrogate (features+GEMMs+apply)", 6: "k_chem_screen",
          8: "k_chem_retry"}

# GRI-3.0 reaction census (chem.inp: 325 reactions, 29 fall-off of which 26 Troe, 12 plain +M,
# 309 reversible) and the FP64 operation model of one right-hand side, counting each exp / log /
# pow as one operation: per species 40 (thermo, S/R-H/RT, concentration), per reaction 14 (rate
# constant, equilibrium constant, rates of progress, six production-rate updates), per fall-off
# reaction 30 more, per third-body efficiency 2 (about 8 per third-body reaction).
def rhs_flops(kk=53, ii=325, nfall=29, ntb=41):
    """FP64 operations of one chemistry right-hand side (CKWYR + the energy equation) with the
    exponentials at the cost of their evaluation (fm_exp: Cody-Waite reduction, 13-term Horner
    series, reconstruction ~45 operations; fm_log ~50): per species 60 (NASA-7 thermo, S/R - H/RT,
    concentrations, production-rate sum), per reaction 14 (concentration products, rate assembly)
    plus one exponential for the Arrhenius factor and one for the equilibrium constant, per fall-off
    reaction ~250 (low-pressure Arrhenius exponential, log10 of Pr, the Troe centre's exponentials,
    its log10 and the blending exp10), per third-body efficiency 2 (8 per third-body row)."""
    return 60 * kk + 14 * ii + 2 * 45 * ii + 250 * nfall + 2 * 8 * ntb


def sdirk_flops(st, kk=53, ii=325):
    """FP64 operations of the implicit integrator on the cells whose counters are the rows of st
    (steps, rejects, newton): per attempt (step or reject) a Jacobian -- two right-hand sides, the
    analytic scatter of up to 36 entries per reaction, the temperature row and the rescaling -- an
    LU of the (kk+1)-square Newton matrix and the error solve; per Newton iteration a right-hand side,
    a forward and a back substitution and the norm."""
    n = kk + 1
    rhs = rhs_flops()
    jac = 2 * rhs + ii * (6 * 4 + 36 * 2) + 3 * n * n
    lu = 2 * n ** 3 // 3 + n * n
    solve = 2 * n * n
    st = np.asarray(st, dtype=np.float64).reshape(-1, 4)
    attempts = st[:, 0] + st[:, 1]
    return float(np.sum(attempts * (jac + lu + solve) + st[:, 2] * (rhs + solve + 2 * n)))


def kernel_table(s):
    out = {}
    for i, name in TIMERS.items():
        ms, n = s.timing(i)
        if n:
            out[name] = {"avg_ms": ms / n, "launches": n, "total_ms": ms}
    return out


def fluid_bytes(nequ, ntot, nblk, ncell):
    """Algorithmic HBM bytes of one k_flux launch: the primitive state once (nequ doubles per cell)
    in, one flux vector per interface (ncell + 1 per block) out."""
    return 8.0 * nequ * (ntot + nblk * (ncell + 1))


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE / WRITE_SIZE summary of this
    workload (profiles/<workload>_hbm_traffic.json, two rocprofv3 --pmc runs scaled by a copy of
    known size), or None."""
    path = os.path.join(REPO, "profiles", TRAFFIC_PROFILES.get(workload, ""))
    if not os.path.isfile(path):
        return None, None
    kernels = json.load(open(path))["kernels"]
    short = kernel.split(" ")[0] + "<"
    # every instantiation launched for the call (the integrator launches its one-wave and its
    # four-wave kernel each time; the one whose mode does not match the list returns at once)
    hits = [rec["read_bytes"] + rec["write_bytes"] for name, rec in kernels.items() if short in name]
    if not hits:
        return None, None
    return sum(hits), os.path.relpath(path, REPO)


TRAFFIC_PROFILES = {"inert1m": "fluid_inert1m_hbm_traffic.json",
                    "inert65536": "fluid_inert65536_hbm_traffic.json",
                    "coupled65536": "coupled65536_hbm_traffic.json"}


def roofline_for(kern, mech, ntot, nblk, ncell, screen_cells, workload="", sdirk_work=None):
    """Roofline of the kernel with the largest total time in the timed pass, with the same record
    for the kernel of the second-largest time under "next"."""
    def one(k):
        roof = _roofline_for(k, mech, ntot, nblk, ncell, screen_cells, sdirk_work)
        traffic, src = measured_traffic(workload, roof["kernel"])
        if traffic is not None:
            roof["traffic"] = traffic
            roof["traffic_source"] = src
        return roof
    roof = one(kern)
    rest = {k: v for k, v in kern.items() if k != roof["kernel"]}
    if rest:
        roof["next"] = one(rest)
    return roof


def _roofline_for(kern, mech, ntot, nblk, ncell, screen_cells, sdirk_work=None):
    name, rec = max(kern.items(), key=lambda kv: kv[1]["total_ms"])
    avg_s = rec["avg_ms"] / 1e3
    if name == "k_flux":
        b = fluid_bytes(mech.nequ, ntot, nblk, ncell)
        a = b / avg_s / 1e9
        return {"bound": "hbm", "kernel": name, "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": a / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": rec["avg_ms"],
                "algorithmic_bytes_per_launch": b}
    if name == "k_stage_end":
        # RK stage: Q0, Qin and the fluxes in, Qout and the primitives out (all cells and ghosts)
        b = 8.0 * mech.nequ * (5 * ntot)
        a = b / avg_s / 1e9
        return {"bound": "hbm", "kernel": name, "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": a / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": rec["avg_ms"],
                "algorithmic_bytes_per_launch": b}
    if name == "k_chem (SDIRK)" and sdirk_work is not None:
        # implicit integrator: the operations its counters say it did, averaged over the timed launches
        a = sdirk_work / avg_s / 1e12
        return {"bound": "valu-fp64", "kernel": name, "achieved": a, "peak": FP64_PEAK_TF,
                "unit": "TFLOP/s", "frac": a / FP64_PEAK_TF, "traffic": None,
                "avg_launch_ms": rec["avg_ms"], "algorithmic_flops_per_launch": sdirk_work,
                "work_model": "per attempt: Jacobian (2 right-hand sides + analytic scatter), LU, error "
                              "solve; per Newton iteration: right-hand side + 2 triangular solves; "
                              "exp/log/pow count as one operation"}
    # chemistry kernels: FP64 vector arithmetic (Heun screen = 2 right-hand sides per cell)
    f = 2.0 * rhs_flops() * screen_cells
    a = f / avg_s / 1e12
    return {"bound": "valu-fp64", "kernel": name, "achieved": a, "peak": FP64_PEAK_TF,
            "unit": "TFLOP/s", "frac": a / FP64_PEAK_TF, "traffic": None, "avg_launch_ms": rec["avg_ms"],
            "algorithmic_flops_per_launch": f}


def oracle_coupled_baseline(cells, nblk, ncell, dx, cfl, steps, strang):
    """The CPU restatement (oracle, gcc -O2, OpenMP over interfaces and cells on the host's cores) on
    `steps` coupled steps of the same domain, split like the timed GPU steps, with the heat-release
    diagnostics.  Returns (seconds, threads)."""
    from oracle.pyoracle import Domain, Oracle
    orc = Oracle()
    d = Domain(orc, cells, nblk, ncell, dx, flag_acu=2, flag_weno=2, select_riemann=2,
               select_time=1)
    t0 = time.perf_counter()
    for _ in range(steps):
        d.step(cfl, react=True, strang=strang, diagnostics=True)
    return time.perf_counter() - t0, int(orc.lib.fo_threads())


def bench_coupled(args, mech, nblk, ncell, dx, nhot, names):
    import torch
    from fortrancfd_amd.lib import Solver
    from fortrancfd_amd.surrogate import load_surrogate
    ntot = nblk * ncell
    cells = ch4_air_cells(names, ntot, nhot, SEED)
    cfl = 0.5
    # surrogate answers verified against the FP64 right-hand side at the integrator's tolerance
    sur = load_surrogate(mech, verify_tol=1.0)
    # the acceptance test alone (training range, element residual, energy-consistent T)
    sur_u = load_surrogate(mech, verify_tol=1e30)

    strang = not args.lie
    calib_copy(args)

    def run(surrogate):
        s = Solver(mech, nblk, ncell, dx, flag_acu=2, flag_weno=2, select_riemann=2, select_time=1,
                   do_reaction=1)
        s.set_step(strang=strang, hrr_each_step=True)
        s.init_cells(cells)
        if surrogate is not None:
            s.set_surrogate(surrogate)
        t = s.advance(cfl, args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t = s.advance(cfl, args.steps, t)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        return s, el, t

    lv = None
    maj = [3 + names.index(k) for k in MAJOR]

    def errors(qf_a, qf_b):
        inner = (slice(None), slice(lv, lv + ncell))
        Ta, Tb = qf_a[inner][..., 2], qf_b[inner][..., 2]
        Ya, Yb = qf_a[inner][..., maj], qf_b[inner][..., maj]
        return {"T_rel_max": float(np.max(np.abs(Ta - Tb) / Tb)),
                "major_species_abs_max": float(np.max(np.abs(Ya - Yb))),
                "frac_cells_T_within_1e-3": float(np.mean(np.abs(Ta - Tb) / Tb < 1e-3)),
                "frac_cells_major_within_1e-4": float(np.mean(np.abs(Ya - Yb).max(axis=-1) < 1e-4))}

    s_sur, el_sur, t_sur = run(sur)
    lv = s_sur.lv
    _, qf_sur = s_sur.download()
    route, nfb = s_sur.source_accepted()
    s_fp, el_fp, t_fp = run(None)
    _, qf_fp = s_fp.download()
    route_fp, _ = s_fp.source_accepted()
    st_fp = s_fp.chem_stats()[route_fp == 0]
    st_fp = st_fp[st_fp[:, 2] > 0]          # cells with Newton work (implicit steps)
    fp64_stiff = {"cells": int(st_fp.shape[0])}
    if st_fp.shape[0]:
        fp64_stiff.update({"steps_mean": float(st_fp[:, 0].mean()), "steps_max": int(st_fp[:, 0].max()),
                           "rejects_mean": float(st_fp[:, 1].mean()), "rejects_max": int(st_fp[:, 1].max()),
                           "newton_mean": float(st_fp[:, 2].mean()), "newton_max": int(st_fp[:, 2].max())})
    err_sur = errors(qf_sur, qf_fp)
    t_err, y_err = err_sur["T_rel_max"], err_sur["major_species_abs_max"]
    s_u, el_u, t_u = run(sur_u)
    _, qf_u = s_u.download()
    route_u, _ = s_u.source_accepted()
    err_u = errors(qf_u, qf_fp)
    del s_u, s_fp

    # per-kernel times: a further pass of the surrogate run with HIP events around every launch,
    # one step at a time so that the implicit integrator's counters of every launch are read
    kt = min(args.steps, 200)
    s_sur.set_timing(True)
    t = t_sur
    sdirk_total = 0.0
    for _ in range(kt):
        t = s_sur.advance(cfl, 1, t)
        route_k, _ = s_sur.source_accepted()
        st_k = s_sur.chem_stats()[route_k == 0]
        sdirk_total += sdirk_flops(st_k[st_k[:, 2] > 0], mech.kk, mech.ii)   # cells with Newton work
    kern = kernel_table(s_sur)
    s_sur.set_timing(False)
    screen_cells = ntot
    roof = roofline_for(kern, mech, ntot, nblk, ncell, screen_cells, args.workload,
                        sdirk_work=sdirk_total / kt)

    cpu = None
    if not args.no_cpu:
        secs, threads = oracle_coupled_baseline(cells, nblk, ncell, dx, cfl, args.cpu_steps, strang)
        cpu = {"value": ntot * args.cpu_steps / secs, "unit": "cell-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"the first {args.cpu_steps} coupled steps of the same {ntot}-cell domain, "
                         f"{'Strang' if strang else 'Lie'}-split with HRR, by the C restatement (gcc "
                         f"-O2, OpenMP, {threads} threads; the host reports {os.cpu_count()} logical "
                         f"CPUs; the reference ships no chemistry integrator): {secs:.2f} s"}
    value = ntot * args.steps / el_sur
    fp64_value = ntot * args.steps / el_fp
    return {
        "metric": "cell-steps/s, coupled 65536-cell GRI-3.0 reacting flow, surrogate + FP64 "
                  "fallback (vs plain FP64 integrator, at stated error)",
        "value": value,
        "unit": "cell-steps/s",
        "ms_per_step": el_sur / args.steps * 1e3,
        "fp64": {"value": fp64_value, "ms_per_step": el_fp / args.steps * 1e3, "sim_time_s": t_fp,
                 "implicit_cells_last_step": fp64_stiff},
        "speedup_vs_fp64": value / fp64_value,
        "error_vs_fp64": dict(err_sur, major_species=list(MAJOR), after_steps=args.warmup + args.steps),
        "chemistry_routes_last_step": {"explicit_screen": int((route == 2).sum()),
                                       "explicit_half_steps": int((route == 3).sum()),
                                       "surrogate": int((route == 1).sum()),
                                       "fp64_implicit": int((route == 0).sum())},
        "unverified_surrogate": {
            "note": "acceptance test without the FP64 verification: training range, element "
                    "residual <= 1e-5, network T within 2e-3 of the energy T",
            "value": ntot * args.steps / el_u, "ms_per_step": el_u / args.steps * 1e3,
            "speedup_vs_fp64": (ntot * args.steps / el_u) / fp64_value,
            "error_vs_fp64": err_u,
            "chemistry_routes_last_step": {"explicit_screen": int((route_u == 2).sum()),
                                           "explicit_half_steps": int((route_u == 3).sum()),
                                           "surrogate": int((route_u == 1).sum()),
                                           "fp64_implicit": int((route_u == 0).sum())}},
        "sim_time_s": t_sur,
        "config": {"workload": f"coupled65536: {nblk} blocks x {ncell} cells, dx {dx} m, GRI-3.0 "
                               f"CH4/air 1 atm 300 K, {nhot} hot spots 1800-2200 K, WENO5-JS+HLLC+"
                               f"SSP-RK3 as a HIP graph, CFL {cfl}, walls, "
                               f"{'Strang' if strang else 'Lie'}-split chemistry with HRR/HRR_dot "
                               f"every step (atol 1e-13, rtol 1e-5)",
                   "cells": ntot, "species": mech.kk, "reactions": mech.ii,
                   "surrogate": "56-4x256 GELU-54 BF16/FP32 MLP, "
                                f"{sur.meta.get('pairs', '?')} training pairs"},
        "kernels": kern,
        "roofline": roof,
        "cpu_baseline": cpu,
    }


def calib_copy(args):
    """A device-to-device copy of known size: the reference point that scales the hardware byte
    counters (FETCH_SIZE / WRITE_SIZE) of a profiled run to bytes."""
    if args.calib_copy_mb:
        import torch
        src = torch.empty(args.calib_copy_mb << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()


def bench_inert(args, mech, nblk, ncell, dx, nhot, names):
    import torch
    from fortrancfd_amd.lib import Solver
    ntot = nblk * ncell
    cells = ch4_air_cells(names, ntot, nhot, SEED)
    s = Solver(mech, nblk, ncell, dx, flag_acu=2, flag_weno=2, select_riemann=2, select_time=1)
    s.init_cells(cells)
    cfl = 0.5
    calib_copy(args)
    t = s.advance(cfl, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t = s.advance(cfl, args.steps, t)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s.set_timing(True)
    s.advance(cfl, min(args.steps, 200), t)
    kern = kernel_table(s)
    s.set_timing(False)
    roof = roofline_for({k: v for k, v in kern.items() if k == "k_flux"}, mech, ntot, nblk, ncell, 0,
                        args.workload)
    cpu = None
    if not args.no_cpu:
        secs = reference_cpu_baseline(cells, nblk, ncell, dx, args.cpu_steps, cfl)
        if secs is not None:
            cpu = {"value": ntot * args.cpu_steps / secs, "unit": "cell-steps/s", "cores": 1,
                   "kind": "reference",
                   "sample": f"{args.cpu_steps} steps of the same {ntot}-cell domain on the "
                             f"reference Fortran (amdflang -O2, one core): {secs:.2f} s"}
    return {
        "metric": "cell-steps/s (GRI-3.0 WENO5+HLLC+SSP-RK3 hydrodynamics, no chemistry)",
        "value": ntot * args.steps / elapsed,
        "unit": "cell-steps/s",
        "ms_per_step": elapsed / args.steps * 1e3,
        "sim_time_s": t,
        "config": {"workload": f"{args.workload}: {nblk} blocks x {ncell} cells, GRI-3.0 CH4/air "
                               f"1 atm 300 K, {nhot} hot spots, WENO5-JS+HLLC+SSP-RK3, CFL {cfl}, "
                               f"walls", "cells": ntot, "species": mech.kk},
        "kernels": kern,
        "roofline": roof,
        "cpu_baseline": cpu,
    }


def bench_chem(args, mech, surrogate):
    """Batched chemistry on 2^20 independent cells (T ~ U(1000, 2600) K, p ~ log-U(1, 50) atm,
    dt ~ log-U(1e-8, 1e-6) s): FP64 integrator, or surrogate + fallback with errors vs FP64."""
    import torch
    from fortrancfd_amd import workloads
    from fortrancfd_amd.surrogate import load_surrogate
    n = args.cells
    seed = workloads.SEED + (5 if surrogate else 3)
    rho, dt, T, Y = workloads.chemistry_batch(mech, n, seed=seed)
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    tr, tdt = d(rho), d(dt)
    Tf, Yf = d(T), d(Y.T)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mech.chem_batch_device(tr, tdt, Tf, Yf)
    torch.cuda.synchronize()
    el_fp = time.perf_counter() - t0
    line = {"unit": "chemistry updates/s", "cells": n}
    if surrogate:
        maj = [names_index(mech, k) for k in MAJOR]

        def infer(verify_tol):
            sur = load_surrogate(mech, verify_tol=verify_tol)
            Ts, Ys = d(T), d(Y.T)
            route = torch.zeros(n, dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            nf = sur.infer_batch(tr, tdt, Ts, Ys, accepted=route)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            r = route.cpu().numpy()
            te = (torch.abs(Ts - Tf) / Tf).cpu().numpy()
            ye = torch.abs(Ys - Yf)[maj].max(dim=0).values.cpu().numpy()
            q = lambda v: {p: float(np.quantile(v, p)) for p in (0.5, 0.9, 0.99, 1.0)}
            return el, {"value": n / el, "speedup_vs_fp64": el_fp / el,
                        "routes": {"explicit_screen": int((r == 2).sum()),
                                   "explicit_half_steps": int((r == 3).sum()),
                                   "surrogate": int((r == 1).sum()), "fp64_implicit": int(nf)},
                        "T_relerr_quantiles": q(te), "major_abserr_quantiles": q(ye),
                        "frac_T_within_1e-3": float((te < 1e-3).mean()),
                        "frac_major_within_1e-4": float((ye < 1e-4).mean())}

        el_u, unver = infer(1e30)
        el_v, ver = infer(1.0)
        line.update({"metric": "chemistry updates/s, surrogate + FP64 fallback (acceptance test "
                               "without FP64 verification), 2^20 fresh cells",
                     "value": n / el_u, "fp64_value": n / el_fp, "speedup_vs_fp64": el_fp / el_u,
                     "unverified": unver, "verified_at_integrator_tolerance": ver})
    else:
        line.update({"metric": "chemistry updates/s, FP64 implicit integrator, 2^20 cells",
                     "value": n / el_fp})
    line["config"] = {"workload": f"{'surrogate1m' if surrogate else 'chem1m'}: {n} independent "
                                  "GRI-3.0 cells, T 1000-2600 K, p 1-50 atm, dt 1e-8-1e-6 s"}
    return line


def names_index(mech, k):
    return mech.names.index(k)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="coupled65536",
                    choices=sorted(WORKLOADS) + ["chem1m", "surrogate1m"])
    ap.add_argument("--cells", type=int, default=1 << 20)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--lie", action="store_true",
                    help="coupled workload: the reference driver's Lie splitting instead of Strang")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="one device copy of this many MiB before the steps (byte-counter reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every block of the domain lives on one device and this "
              "build has no multi-GPU path (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    from fortrancfd_amd.lib import GRI30, Mechanism

    mech = Mechanism()
    names = species_names(GRI30)
    if args.workload in ("chem1m", "surrogate1m"):
        line = bench_chem(args, mech, args.workload == "surrogate1m")
        steps = 1
    else:
        nblk, ncell, dx, nhot, react = WORKLOADS[args.workload]
        fn = bench_coupled if react else bench_inert
        line = fn(args, mech, nblk, ncell, dx, nhot, names)
        steps = args.steps
    out = {"metric": line.pop("metric"), "value": line.pop("value"), "unit": line.pop("unit"),
           "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": line.pop("ms_per_step", None), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    out.update(line)
    prop = torch.cuda.get_device_properties(0)
    out["device"] = {"name": prop.name, "compute_units": prop.multi_processor_count,
                     "arch": getattr(prop, "gcnArchName", "")}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
