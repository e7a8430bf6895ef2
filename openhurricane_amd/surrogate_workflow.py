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


def make_context(mech, device_pointers=True, jacobian=1):
    from . import _lib
    from .chemistry import ChemistrySource

    o = _lib.default_opts(atol=1e-10, rtol=1e-6, atol_inc=1e-7, rtol_inc=1e-3, jacobian=jacobian,
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


def error_stats(e, kind_t=None):
    import torch

    ef = e.float()
    sub = ef[:: max(1, ef.numel() // 1000000)]
    qs = torch.quantile(sub, torch.tensor([0.5, 0.99, 0.999], device=e.device)).tolist()
    out = {"max": float(e.max()), "p999": qs[2], "p99": qs[1], "p50": qs[0],
           "mean": float(e.mean()), "frac_le_1e-3": float((e <= 1e-3).float().mean())}
    if kind_t is not None:
        for k, name in ((0, "front"), (1, "cold"), (2, "burnt")):
            sel = kind_t == k
            if bool(sel.any()):
                out[f"max_{name}"] = float(e[sel].max())
    return out


def _timed(fn, steps):
    import torch

    t = 0.0
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
    return t, r


def evaluate(mech, cs, nets, n, seed, front, configs, dt=1e-6, steps=2):
    """Plain integrator vs gated surrogate + fallback on n held-out cells.

    nets: one or two trained surrogates; configs: dicts with gate_tol,
    peer_tol (None: single network) and precision (0 BF16, 1 FP8)."""
    import torch

    dev = torch.device("cuda:0")
    burner = make_context(mech, device_pointers=False)
    g, kind = states_on_device(mech, n, seed, front, burner, dev)
    kind_t = torch.from_numpy(kind).to(dev)
    dtv = torch.full((n,), dt, dtype=torch.float64, device=dev)
    # plain FP64 integrator: reference result and timing (the last of `steps`
    # flow steps, with the carried sub-step memory and cell schedule)
    sub = torch.zeros(n, dtype=torch.float64, device=dev)
    res = {}

    def plain():
        T, Y = g["T"].clone(), g["Y"].clone()
        _, st = cs.solve_update(T, g["p"], g["rho"], Y, dtv, 1.0, subDt=sub)
        res["T"], res["Y"], res["st"] = T, Y, st

    t_plain, _ = _timed(plain, steps)
    Tr, Yr = res["T"], res["Y"]
    out = {"cells": n, "plain": {"seconds": t_plain, "cells_per_s": n / t_plain,
                                 "newton_iters": res["st"]["newton_iters"]},
           "plain_by_kind_seconds": {}, "gated": []}
    for k, name in ((0, "front"), (1, "cold"), (2, "burnt")):
        idx = torch.nonzero(kind_t == k).flatten()
        if idx.numel() == 0:
            continue
        sl = {key: (v[:, idx] if v.dim() == 2 else v[idx]).contiguous() for key, v in g.items()}
        subk = torch.zeros(idx.numel(), dtype=torch.float64, device=dev)
        dk = dtv[idx].contiguous()

        def plain_k():
            T, Y = sl["T"].clone(), sl["Y"].clone()
            cs.solve_update(T, sl["p"], sl["rho"], Y, dk, 1.0, subDt=subk)

        out["plain_by_kind_seconds"][name] = _timed(plain_k, steps)[0]
    # network alone (inference kernel, no gate, no fallback)
    for prec in (0, 1):
        nets[0].set_precision(prec)
        t_inf, _ = _timed(lambda: nets[0].predict(g["T"], g["p"], g["Y"]), 3)
        out[f"inference_{'fp8' if prec else 'bf16'}"] = {"seconds": t_inf, "cells_per_s": n / t_inf}
    for cfg in configs:
        lead = nets[0]
        lead.set_precision(cfg.get("precision", 0))
        if len(nets) > 1:
            nets[1].set_precision(cfg.get("precision", 0))
        pt = cfg.get("peer_tol")
        lead.set_peer(nets[1] if (pt is not None and len(nets) > 1) else None, pt or 0.0)
        subs = torch.zeros(n, dtype=torch.float64, device=dev)

        def gated():
            T, Y = g["T"].clone(), g["Y"].clone()
            nf, st = lead.solve_update(T, g["p"], g["rho"], Y, dtv, 1.0, subDt=subs,
                                       gate_tol=cfg["gate_tol"])
            res["gT"], res["gY"], res["nf"] = T, Y, nf

        t, _ = _timed(gated, steps)
        eT, eY = errors(res["gT"], res["gY"], Tr, Yr)
        e = torch.maximum(eT, eY)
        # a cell that fell back carries the plain integrator's result bit for
        # bit, an accepted one the network's: acceptance per cell kind
        acc = (res["gT"] != Tr) | (res["gY"] != Yr).any(0)
        by_kind = {}
        for k, name in ((0, "front"), (1, "cold"), (2, "burnt")):
            sel = kind_t == k
            if bool(sel.any()):
                by_kind[name] = float(acc[sel].float().mean())
        r = dict(cfg)
        r.update({"seconds": t, "cells_per_s": n / t, "speedup": t_plain / t,
                  "fallback_frac": res["nf"] / n, "accepted_by_kind": by_kind,
                  "err": error_stats(e, kind_t),
                  "max_rel_err_T": float(eT.max()), "max_rel_err_Ymajor": float(eY.max())})
        out["gated"].append(r)
        log(json.dumps(r))
    lead.set_peer(None, 0.0)
    lead.set_precision(0)
    return out


def run(argv=None):
    """Generate pairs, train, evaluate; returns the result dictionary."""
    ap = argparse.ArgumentParser()
    ap.add_argument("--mech", default="syn32", choices=["syn32", "syn54"])
    ap.add_argument("--pairs", type=int, default=1 << 24)
    ap.add_argument("--eval", type=int, default=1 << 22)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--nets", type=int, default=2, choices=[1, 2])
    ap.add_argument("--delta-T", type=float, default=1e-2, help="target scale of dT [K]")
    ap.add_argument("--delta-Y", type=float, default=1e-4, help="target scale of d ln Y")
    ap.add_argument("--plain-targets", action="store_true", help="train on plain increments")
    ap.add_argument("--peer-tol", type=float, nargs="+", default=[1e-5, 3e-5, 1e-4, 3e-4])
    ap.add_argument("--gate", type=float, default=1e-2, help="element-mass gate")
    ap.add_argument("--save", default=None)
    ap.add_argument("--jacobian", type=int, default=1,
                    help="Jacobian of the plain integrator (0: the reference's dKp/dT sign "
                         "convention, 1: exact)")
    args = ap.parse_args(argv)

    import torch

    from .chemistry import Mechanism
    from .surrogate import Surrogate

    ns, nr, seed, width, front = {"syn32": (32, 150, 1, 256, 0.2),
                                  "syn54": (54, 270, 2, 512, 0.3)}[args.mech]
    mech = Mechanism.synthetic(ns, nr, seed=seed)
    cs = make_context(mech, jacobian=args.jacobian)
    nets = [Surrogate(cs, width=width, depth=3, seed=k) for k in range(args.nets)]
    delta = None
    if not args.plain_targets:
        delta = np.full(ns + 1, args.delta_Y, dtype=np.float32)
        delta[0] = args.delta_T
    for s in nets:
        s.set_target_scale(delta)
    t0 = time.perf_counter()
    X, Yt = generate_pairs(mech, cs, nets[0], args.pairs, seed=1000, front=front)
    t_gen = time.perf_counter() - t0
    hist, t_fit = [], 0.0
    for k, s in enumerate(nets):
        s.fit_norm(X, Yt, 1e-6)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hist.append(s.fit(X, Yt, epochs=args.epochs, batch=args.batch, lr=args.lr,
                          lr_end=args.lr / 100, seed=k, log=log))
        torch.cuda.synchronize()
        t_fit += time.perf_counter() - t0
        if args.save:
            s.save(f"{args.save}.{k}")
    del X, Yt
    configs = [{"gate_tol": 1e30, "peer_tol": None}, {"gate_tol": args.gate, "peer_tol": None}]
    if len(nets) > 1:
        configs += [{"gate_tol": args.gate, "peer_tol": t} for t in args.peer_tol]
        configs += [{"gate_tol": args.gate, "peer_tol": t, "precision": 1} for t in args.peer_tol]
    res = evaluate(mech, cs, nets, args.eval, seed=777777, front=front, configs=configs)
    res.update({"mechanism": f"synthetic {ns}sp/{nr}rxn", "pairs": args.pairs,
                "epochs": args.epochs, "batch": args.batch, "width": width, "depth": 3,
                "nets": len(nets), "jacobian": args.jacobian, "target_scale": None if delta is None else
                {"dT": args.delta_T, "dlnY": args.delta_Y},
                "loss_history": hist, "seconds_generate": t_gen, "seconds_train": t_fit,
                "train_samples_per_s": len(nets) * args.pairs * args.epochs / t_fit})
    return res


def main(argv=None):
    print(json.dumps(run(argv)), flush=True)


if __name__ == "__main__":
    main()
