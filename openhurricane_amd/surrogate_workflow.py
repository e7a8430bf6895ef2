"""Surrogate workflow on the synthetic-mechanism workloads: generate state
pairs with the plain GPU integrator, train the MLP on the GPU with the
library's own BF16 kernels, and evaluate gated surrogate + fallback against
the plain integrator on held-out cells.

    python -m openhurricane_amd.surrogate_workflow --pairs 16777216 --eval 4194304

Error metric (per held-out cell, against the plain FP64 integrator, which
equals the FP64 oracle to rounding): relative error of T and of every
species whose reference mass fraction is >= 1e-3 ("major" species).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

MAJOR_Y = 1e-3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_context(mech, device_pointers=True):
    from . import _lib
    from .chemistry import ChemistrySource

    o = _lib.default_opts(atol=1e-10, rtol=1e-6, atol_inc=1e-7, rtol_inc=1e-3, jacobian=1,
                          device_pointers=int(device_pointers), kernel=2, schedule=1)
    return ChemistrySource(mech, opts=o)


def states_on_device(mech, n, seed, front, burner, dev):
    import torch

    from .sampler import synthetic_workload

    t = mech.tables()

    def burn(T, p, rho, Y, dt):
        burner.solve_update(T, p, rho, Y, dt, 1.0)

    st = synthetic_workload(mech.species, t["W"], n, seed=seed, front=front, burn=burn)
    kind = st.pop("kind")
    g = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in st.items()}
    return g, kind


def generate_pairs(mech, cs, sur, n, seed, front=0.2, chunk=1 << 21, dt=1e-6):
    """Raw features / targets of n pairs (device float32), in chunks."""
    import torch

    dev = torch.device("cuda:0")
    burner = make_context(mech, device_pointers=False)
    Xs, Ys = [], []
    done = 0
    while done < n:
        m = min(chunk, n - done)
        g, _ = states_on_device(mech, m, seed + done, front, burner, dev)
        T1, Y1 = g["T"].clone(), g["Y"].clone()
        dtv = torch.full((m,), dt, dtype=torch.float64, device=dev)
        cs.solve_update(T1, g["p"], g["rho"], Y1, dtv, 1.0)
        X, Yt = sur.features(g["T"], g["p"], g["Y"], T1, Y1)
        Xs.append(X)
        Ys.append(Yt)
        done += m
        log(f"pairs: {done}/{n}")
    return torch.cat(Xs), torch.cat(Ys)


def errors(T, Y, Tr, Yr):
    """Per-cell max relative error of T and of the major species."""
    import torch

    eT = (T - Tr).abs() / Tr
    major = Yr >= MAJOR_Y
    eY = torch.where(major, (Y - Yr).abs() / Yr.clamp_min(1e-300), torch.zeros_like(Yr)).amax(0)
    return eT, eY


def evaluate(mech, cs, sur, n, seed, front, gate_tols, dt=1e-6, steps=2):
    """Plain integrator vs gated surrogate + fallback on n held-out cells."""
    import torch

    dev = torch.device("cuda:0")
    burner = make_context(mech, device_pointers=False)
    g, kind = states_on_device(mech, n, seed, front, burner, dev)
    dtv = torch.full((n,), dt, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    cs.set_stream(stream.cuda_stream)
    # plain FP64 integrator: reference result and timing (the second of two
    # flow steps, with the carried sub-step memory and cell schedule)
    sub = torch.zeros(n, dtype=torch.float64, device=dev)
    times = []
    for _ in range(steps):
        T, Y = g["T"].clone(), g["Y"].clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, st_plain = cs.solve_update(T, g["p"], g["rho"], Y, dtv, 1.0, subDt=sub)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    t_plain = times[-1]
    Tr, Yr = T, Y
    out = {"cells": n, "plain": {"seconds": t_plain, "cells_per_s": n / t_plain,
                                 "newton_iters": st_plain["newton_iters"]}, "gated": []}
    for tol in gate_tols:
        subs = torch.zeros(n, dtype=torch.float64, device=dev)
        times = []
        for _ in range(steps):
            T, Y = g["T"].clone(), g["Y"].clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            nf, st = sur.solve_update(T, g["p"], g["rho"], Y, dtv, 1.0, subDt=subs, gate_tol=tol)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        eT, eY = errors(T, Y, Tr, Yr)
        e = torch.maximum(eT, eY)
        res = {"gate_tol": tol, "seconds": times[-1], "cells_per_s": n / times[-1],
               "speedup": t_plain / times[-1], "fallback": nf, "fallback_frac": nf / n,
               "max_rel_err_T": float(eT.max()), "max_rel_err_Ymajor": float(eY.max()),
               "mean_rel_err": float(e.mean()), "frac_cells_err_le_1e-3": float((e <= 1e-3).float().mean())}
        for k, name in ((0, "front"), (1, "cold"), (2, "burnt")):
            sel = torch.from_numpy(kind == k).to(dev)
            if bool(sel.any()):
                res[f"max_err_{name}"] = float(e[sel].max())
        out["gated"].append(res)
        log(json.dumps(res))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--mech", default="syn32", choices=["syn32", "syn54"])
    ap.add_argument("--pairs", type=int, default=1 << 24)
    ap.add_argument("--eval", type=int, default=1 << 22)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--gate", type=float, nargs="+", default=[1e-4, 1e-3, 1e30])
    ap.add_argument("--save", default=None)
    args = ap.parse_args(argv)

    import torch

    from .chemistry import Mechanism
    from .surrogate import Surrogate

    ns, nr, seed, width, front = {"syn32": (32, 150, 1, 256, 0.2),
                                  "syn54": (54, 270, 2, 512, 0.3)}[args.mech]
    mech = Mechanism.synthetic(ns, nr, seed=seed)
    cs = make_context(mech)
    sur = Surrogate(cs, width=width, depth=3, seed=0)
    t0 = time.perf_counter()
    X, Yt = generate_pairs(mech, cs, sur, args.pairs, seed=1000, front=front)
    t_gen = time.perf_counter() - t0
    sur.fit_norm(X, Yt, 1e-6)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = sur.fit(X, Yt, epochs=args.epochs, batch=args.batch, lr=args.lr, lr_end=args.lr / 100,
                   log=log)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    if args.save:
        sur.save(args.save)
    del X, Yt
    res = evaluate(mech, cs, sur, args.eval, seed=777777, front=front, gate_tols=args.gate)
    res.update({"mechanism": f"synthetic {ns}sp/{nr}rxn", "pairs": args.pairs,
                "epochs": args.epochs, "batch": args.batch, "width": width, "depth": 3,
                "loss_history": hist, "seconds_generate": t_gen, "seconds_train": t_fit,
                "train_samples_per_s": args.pairs * args.epochs / t_fit})
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
