"""Training and loading of the chemistry surrogate.

    python -m fortrancfd_amd.surrogate train [--pairs N] [--epochs E] [--batch 8192] [--out PATH]
    python -m fortrancfd_amd.surrogate eval [--cells N] [--weights PATH]

Training pairs come from the batched-chemistry generator (workloads.chemistry_batch: T ~ U(1000,
2600) K, p ~ log-U(1, 50) atm, dt ~ log-U(1e-8, 1e-6) s, mixing-line compositions) solved by this
package's FP64 integrator on the GPU, optionally joined by the stiff-cell (state, result) pairs of
coupled reacting runs with other hot-spot seeds than the benchmark's (ignition kernels of pristine
fresh mixture, which the mixing line never produces); --coupled-strang records those runs' source
calls over dt/2 as the benchmark's Strang splitting makes them (below the batch generator's smallest
dt), instead of over dt (Lie).  The network is trained on the GPU by the library's own
forward / backward / Adam kernels (fcfd_sur_train_step): Adam, learning rate 1e-3 on a cosine
schedule, batch 8192.  Species enter and leave the network through the Box-Cox transform
b(Y) = 10 (Y^0.1 - 1); inputs are z-scored with the training set's mean and standard deviation;
the training range used by the fallback test is the training set's per-feature [min, max]
widened by 2 % of its width on each side.

Device memory comes from torch (plumbing); every computation runs in libfcfd_hip.so.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import workloads
from .lib import DATA_DIR, Mechanism, Surrogate

DEFAULT_WEIGHTS = os.path.join(DATA_DIR, "surrogate_gri30.npz")
RANGE_MARGIN = 0.02
P_ATM = 101325.0
RU_SI = 8.31451


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the surrogate runs on the GPU: no HIP device is visible")
    return torch


def boxcox(Y):
    """Box-Cox transform of mass fractions, lambda = 1/10 (the network's species coordinate)."""
    return 10.0 * (np.maximum(Y, 0.0) ** 0.1 - 1.0)


def raw_features(mech, rho, dt, T, Y):
    """(n, 56) raw inputs: Box-Cox(Y), T, log(p / 1 atm), log dt (host numpy, FP64)."""
    p = rho * RU_SI * T * (Y / mech.weights).sum(axis=1) * 1e3
    return np.column_stack([boxcox(Y), T, np.log(p / P_ATM), np.log(dt)])


def solve_pairs(mech, rho, dt, T, Y, chunk=131072, log=print):
    """FP64 integrator results (T1, Y1) for host arrays; progress per chunk."""
    torch = _torch()
    n = T.size
    T1, Y1 = np.empty(n), np.empty_like(Y)
    t0 = time.time()
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
        tr, tdt, tT, tY = d(rho[a:b]), d(dt[a:b]), d(T[a:b]), d(Y[a:b].T)
        mech.chem_batch_device(tr, tdt, tT, tY)
        T1[a:b] = tT.cpu().numpy()
        Y1[a:b] = tY.cpu().numpy().T
        log(f"  solved {b}/{n} pairs ({time.time() - t0:.1f} s)")
    return T1, Y1


def coupled_pairs(mech, seeds, steps, every=2, nblk=16, ncell=4096, dx=25e-6, cfl=0.5, log=print,
                  strang=False):
    """(state, FP64 result) pairs of the cells the explicit screen sends to the implicit integrator
    in coupled reacting runs: the 65536-cell CH4/air domain with three 32-cell hot spots of the
    given seeds, stepped with the FP64 chemistry, split Lie-style (fluid, then a dt source) or
    (strang) as the benchmark splits (dt/2 source, fluid, dt/2 source); every `every`-th step the stiff cells' states
    before and after each source call are recorded (the steps in between replay the captured step).
    Returns host arrays (rho, dt, T0, Y0, T1, Y1)."""
    from .lib import Solver
    out = {k: [] for k in ("rho", "dt", "T0", "Y0", "T1", "Y1")}
    t0 = time.time()
    for seed in seeds:
        s = Solver(mech, nblk, ncell, dx, flag_acu=2, flag_weno=2, select_riemann=2,
                   select_time=1, do_reaction=1)
        s.set_step(strang=strang, hrr_each_step=False)
        s.init_cells(workloads.coupled_cells(mech, nblk, ncell, seed=seed))
        lv = s.lv
        inner = (slice(None), slice(lv, lv + ncell))
        n0 = sum(len(v) for v in out["T0"])
        t = 0.0

        def source(h, tnow):
            q0 = s.download()[1][inner].reshape(-1, mech.nequ)
            s.solve_source(h, tnow, diagnostics=False)
            route, _ = s.source_accepted()
            stiff = route == 0
            if stiff.any():
                q1 = s.download()[1][inner].reshape(-1, mech.nequ)
                y0 = np.maximum(q0[stiff, 3:], 0.0)
                out["rho"].append(q0[stiff, 0])
                out["dt"].append(np.full(stiff.sum(), h))
                out["T0"].append(q0[stiff, 2])
                out["Y0"].append(y0 / y0.sum(axis=1, keepdims=True))
                out["T1"].append(q1[stiff, 2])
                out["Y1"].append(q1[stiff, 3:])
            s.refresh()

        for step in range(0, steps, every):
            if every > 1:
                t = s.advance(cfl, every - 1, t)         # device-resident steps in between
            dt = s.calc_dt(cfl)
            if strang:
                source(0.5 * dt, t)
                s.solve_fluid(dt)
                source(0.5 * dt, t + dt)
            else:
                s.solve_fluid(dt)
                source(dt, t + dt)
            t += dt
            if (step + every) % 1000 < every:
                log(f"    seed {seed}: step {step + every} ({time.time() - t0:.1f} s)")
        n1 = sum(len(v) for v in out["T0"])
        log(f"  coupled run seed {seed}: {steps} steps, {n1 - n0} stiff-cell pairs "
            f"({time.time() - t0:.1f} s)")
    return tuple(np.concatenate(out[k]) for k in ("rho", "dt", "T0", "Y0", "T1", "Y1"))


def fit_norm(mech, feats, dY, dTv):
    kk = mech.kk
    nm = {k: np.zeros(64) for k in Surrogate.NORM_FIELDS}
    nm["in_std"][:] = 1.0
    nm["out_std"][:] = 1.0
    mu, sd = feats.mean(axis=0), feats.std(axis=0)
    lo, hi = feats.min(axis=0), feats.max(axis=0)
    w = hi - lo
    nm["in_mean"][:feats.shape[1]] = mu
    nm["in_std"][:feats.shape[1]] = np.where(sd > 1e-12, sd, 1.0)
    nm["lo"][:feats.shape[1]] = lo - RANGE_MARGIN * w
    nm["hi"][:feats.shape[1]] = hi + RANGE_MARGIN * w
    nm["lo"][feats.shape[1]:] = -1e300
    nm["hi"][feats.shape[1]:] = 1e300
    out = np.column_stack([dY, dTv])
    omu, osd = out.mean(axis=0), out.std(axis=0)
    nm["out_mean"][:kk + 1] = omu
    nm["out_std"][:kk + 1] = np.where(osd > 1e-30, osd, 1.0)
    return nm


def save(path, params, norm, meta):
    np.savez_compressed(path, params=params.astype(np.float32),
                        **{k: np.asarray(norm[k], np.float64) for k in Surrogate.NORM_FIELDS},
                        meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))


def load(path=DEFAULT_WEIGHTS):
    z = np.load(path, allow_pickle=False)
    norm = {k: z[k] for k in Surrogate.NORM_FIELDS}
    meta = json.loads(bytes(z["meta"]).decode()) if "meta" in z else {}
    return z["params"], norm, meta


def load_surrogate(mech, path=DEFAULT_WEIGHTS, resid_tol=1e-5, t_tol=2e-3, verify_tol=1.0):
    params, norm, meta = load(path)
    s = Surrogate(mech)
    s.set_params(params)
    s.set_norm(norm)
    s.set_thresholds(resid_tol, t_tol, verify_tol)
    s.meta = meta
    return s


def train(mech, n_pairs, epochs, batch=8192, lr=1e-3, seed=workloads.SEED + 4, log=print,
          coupled_seeds=(), coupled_steps=5000, coupled_every=2, coupled_strang=False):
    """Generate n_pairs batched-chemistry training pairs (plus, with coupled_seeds, the stiff-cell
    pairs of coupled runs with those hot-spot seeds), fit the scaling and train for `epochs`
    passes.  Returns (Surrogate, params, norm, meta)."""
    torch = _torch()
    t0 = time.time()
    rho, dt, T0, Y0 = workloads.chemistry_batch(mech, n_pairs, seed=seed)
    log(f"generated {n_pairs} states ({time.time() - t0:.1f} s); solving with the FP64 integrator")
    T1, Y1 = solve_pairs(mech, rho, dt, T0, Y0, log=log)
    n_batch = n_pairs
    if coupled_seeds:
        c = coupled_pairs(mech, coupled_seeds, coupled_steps, coupled_every, log=log, strang=coupled_strang)
        rho, dt, T0, Y0, T1, Y1 = (np.concatenate([a, b]) for a, b in zip((rho, dt, T0, Y0, T1, Y1), c))
        n_pairs = T0.size
    t_solve = time.time() - t0
    feats = raw_features(mech, rho, dt, T0, Y0)
    norm = fit_norm(mech, feats, boxcox(Y1) - boxcox(Y0), T1 - T0)
    s = Surrogate(mech)
    s.set_norm(norm)
    s.set_params(he_uniform(seed))
    npad = (n_pairs + 63) // 64 * 64
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    X = torch.zeros((npad, 64), dtype=torch.int16, device="cuda")
    Tg = torch.zeros((npad, 64), dtype=torch.float32, device="cuda")
    tr, tdt, tT0, tY0 = d(rho), d(dt), d(T0), d(Y0.T)
    s.features(tr, tdt, tT0, tY0, X)
    s.targets(tT0, tY0, d(T1), d(Y1.T), Tg)
    del tr, tdt, tT0, tY0
    steps_per_epoch = n_pairs // batch
    total = steps_per_epoch * epochs
    g = torch.Generator(device="cuda").manual_seed(seed)
    hist = []
    step = 0
    t1 = time.time()
    for ep in range(epochs):
        perm = torch.randperm(n_pairs, generator=g, device="cuda").to(torch.int32)
        for b in range(steps_per_epoch):
            lr_t = 0.5 * lr * (1.0 + math.cos(math.pi * step / total))
            want = b == steps_per_epoch - 1
            loss = s.train_step(X, Tg, perm[b * batch:(b + 1) * batch], lr_t, want_loss=want)
            step += 1
        hist.append(loss)
        log(f"  epoch {ep + 1}/{epochs}: loss {loss:.4e} (lr {lr_t:.2e}, {time.time() - t1:.1f} s)")
    torch.cuda.synchronize()
    t_train = time.time() - t1
    meta = {"pairs": n_pairs, "batched_pairs": n_batch, "coupled_pairs": n_pairs - n_batch,
            "coupled_seeds": list(coupled_seeds), "coupled_steps": coupled_steps,
            "coupled_every": coupled_every, "coupled_split": "strang" if coupled_strang else "lie",
            "epochs": epochs, "batch": batch, "lr": lr, "seed": seed,
            "steps": total, "final_loss": float(hist[-1]), "loss_history": [float(v) for v in hist],
            "solve_s": t_solve, "train_s": t_train,
            "train_samples_per_s": total * batch / t_train}
    return s, s.get_params(), norm, meta


def evaluate(mech, sur, n_cells, seed=workloads.SEED + 5, log=print):
    """Surrogate + fallback vs the FP64 integrator on n fresh batched-chemistry cells."""
    torch = _torch()
    rho, dt, T0, Y0 = workloads.chemistry_batch(mech, n_cells, seed=seed)
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    tr, tdt = d(rho), d(dt)
    Ts, Ys = d(T0), d(Y0.T)
    Tf, Yf = Ts.clone(), Ys.clone()
    acc = torch.zeros(n_cells, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.time()
    nf = sur.infer_batch(tr, tdt, Ts, Ys, accepted=acc)
    torch.cuda.synchronize()
    t_sur = time.time() - t0
    t0 = time.time()
    mech.chem_batch_device(tr, tdt, Tf, Yf)
    torch.cuda.synchronize()
    t_fp64 = time.time() - t0
    route = acc.cpu().numpy()
    a = route == 1
    Ts, Tf = Ts.cpu().numpy(), Tf.cpu().numpy()
    Ys, Yf = Ys.cpu().numpy().T, Yf.cpu().numpy().T
    major = [mech.names.index(k) for k in ("CH4", "O2", "N2", "CO2", "H2O", "CO", "H2")]
    terr = np.abs(Ts - Tf) / Tf
    yerr = np.abs(Ys - Yf)[:, major].max(axis=1)
    res = {"cells": n_cells, "fallback": int(nf), "accepted_frac": float(a.mean()),
           "explicit_frac": float((route == 2).mean()),
           "T_relerr_max": float(terr.max()), "T_relerr_mean_accepted": float(terr[a].mean()) if a.any() else 0.0,
           "major_abserr_max": float(yerr.max()),
           "major_abserr_mean_accepted": float(yerr[a].mean()) if a.any() else 0.0,
           "frac_T_within_1e-3": float((terr < 1e-3).mean()),
           "frac_major_within_1e-4": float((yerr < 1e-4).mean()),
           "surrogate_s": t_sur, "fp64_s": t_fp64, "speedup": t_fp64 / t_sur}
    log(json.dumps(res))
    return res


def he_uniform(seed):
    """He-uniform initial weights (zero biases) in the parameter order of fcfd_sur_set_params."""
    rng = np.random.default_rng(seed)
    K = (64, 256, 256, 256, 256)
    N = (256, 256, 256, 256, 64)
    parts = []
    for l, (k, n) in enumerate(zip(K, N)):
        fan_in = 56 if l == 0 else k
        lim = math.sqrt(6.0 / fan_in) if l < 4 else math.sqrt(3.0 / fan_in)
        W = rng.uniform(-lim, lim, (k, n))
        if l == 0:
            W[56:, :] = 0.0
        if l == 4:
            W[:, 54:] = 0.0
        parts += [W.ravel(), np.zeros(n)]
    return np.concatenate(parts).astype(np.float32)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m fortrancfd_amd.surrogate")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--pairs", type=int, default=1 << 22)
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--batch", type=int, default=8192)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    t.add_argument("--eval-cells", type=int, default=65536)
    t.add_argument("--coupled-seeds", type=int, nargs="*", default=[],
                   help="hot-spot seeds of coupled runs whose stiff cells join the training set")
    t.add_argument("--coupled-steps", type=int, default=5000)
    t.add_argument("--coupled-every", type=int, default=2)
    t.add_argument("--coupled-strang", action="store_true",
                   help="record the coupled runs' source calls over dt/2 (Strang) instead of dt (Lie)")
    e = sub.add_parser("eval")
    e.add_argument("--cells", type=int, default=65536)
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    mech = Mechanism()
    if a.cmd == "train":
        s, params, norm, meta = train(mech, a.pairs, a.epochs, a.batch, a.lr, log=log,
                                      coupled_seeds=a.coupled_seeds, coupled_steps=a.coupled_steps,
                                      coupled_every=a.coupled_every, coupled_strang=a.coupled_strang)
        s.set_thresholds()
        if a.eval_cells:
            meta["eval"] = evaluate(mech, s, a.eval_cells, log=log)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        save(a.out, params, norm, meta)
        log(f"wrote {a.out}")
    else:
        evaluate(mech, load_surrogate(mech, a.weights), a.cells, log=log)
    return 0


if __name__ == "__main__":
    sys.exit(main())
