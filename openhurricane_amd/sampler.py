"""Seeded synthetic cell states for the benchmark and parity workloads.

Synthetic-mechanism workloads (32 and 54 species): fresh fuel/air mixtures
(fuel = equimolar H2 + CH4, mixture fraction Z ~ U(0, 0.1), 1e-6 added to
every species mass fraction) at T ~ U(1100, 1600) K form the ignition-front
share; the near-equilibrium share is half cold fresh mixture at
T ~ U(300, 1000) K and half burnt states -- fresh mixtures started at
T ~ U(1800, 2600) K and integrated for 1e-2 s by the caller's integrator,
drawn from a pool and replicated.

H2-air states (the 9-species mechanism): T ~ U(900, 2500) K, p ~ U(0.5, 5) atm,
mixture fraction Z ~ U(0, 1) of pure H2 against air (O2 0.2329 / N2 0.7671 by
mass), 1e-6 added to each radical mass fraction (H, O, OH, HO2, H2O2), then
renormalised; rho = p W / (Ru T) (perfect gas, perfectGas::rhom,
src/speciesTransport/species/src/perfectGas.cpp:36-40).
"""
from __future__ import annotations

import numpy as np

RU = 1.380649e-23 * 6.02214076e23 * 1e3
PATM = 1.01325e5
RADICALS = ("H", "O", "OH", "HO2", "H2O2")


def mixture_density(p, T, Y, W):
    """rho = p / (Ru T sum(Y/W)) for an (ns, n) block Y."""
    return p / (RU * T * (Y / W[:, None]).sum(axis=0))


def h2air_states(species, W, n, seed, T_range=(900.0, 2500.0), p_atm=(0.5, 5.0),
                 radical=1e-6):
    rng = np.random.Generator(np.random.PCG64(seed))
    T = rng.uniform(*T_range, size=n)
    p = rng.uniform(*p_atm, size=n) * PATM
    Z = rng.uniform(0.0, 1.0, size=n)
    idx = {s: i for i, s in enumerate(species)}
    ns = len(species)
    Y = np.zeros((ns, n))
    Y[idx["H2"]] = Z
    Y[idx["O2"]] = (1.0 - Z) * 0.2329
    Y[idx["N2"]] = (1.0 - Z) * 0.7671
    for r in RADICALS:
        if r in idx:
            Y[idx[r]] += radical
    Y /= Y.sum(axis=0, keepdims=True)
    rho = mixture_density(p, T, Y, np.asarray(W))
    return {"T": T, "p": p, "rho": rho, "Y": np.ascontiguousarray(Y)}


def fresh_states(species, W, n, rng, T_range, p_atm=(0.5, 5.0), Z_range=(0.0, 0.1),
                 seed_y=1e-6):
    """Fuel/air mixtures of a synthetic C/H/O/N mechanism (see module docstring)."""
    W = np.asarray(W)
    idx = {s: i for i, s in enumerate(species)}
    ns = len(species)
    T = rng.uniform(*T_range, size=n)
    p = rng.uniform(*p_atm, size=n) * PATM
    Z = rng.uniform(*Z_range, size=n)
    fuel = [f for f in ("H2", "CH4") if f in idx]
    wf = np.array([W[idx[f]] for f in fuel])
    yf = wf / wf.sum()  # equal moles of each fuel species
    Y = np.full((ns, n), seed_y)
    for f, y in zip(fuel, yf):
        Y[idx[f]] += Z * y
    Y[idx["O2"]] += (1.0 - Z) * 0.2329
    Y[idx["N2"]] += (1.0 - Z) * 0.7671
    Y /= Y.sum(axis=0, keepdims=True)
    rho = mixture_density(p, T, Y, W)
    return {"T": T, "p": p, "rho": rho, "Y": np.ascontiguousarray(Y)}


def synthetic_workload(species, W, n, seed, front=0.2, burn=None, pool=4096,
                       burn_time=1e-2):
    """Cells of the synthetic-mechanism workloads.

    front: ignition-front share; the rest is near-equilibrium, half cold and
    half burnt.  burn(T, p, rho, Y, dt) integrates states in place (T and Y
    updated) and is only called on a pool of at most `pool` states, which the
    burnt cells then sample; without it the burnt share is cold as well."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W = np.asarray(W)
    n_front = int(round(front * n))
    n_rest = n - n_front
    n_burnt = n_rest // 2 if burn is not None else 0
    n_cold = n_rest - n_burnt
    parts = [fresh_states(species, W, n_front, rng, (1100.0, 1600.0)),
             fresh_states(species, W, n_cold, rng, (300.0, 1000.0))]
    if n_burnt:
        m = min(pool, n_burnt)
        b = fresh_states(species, W, m, rng, (1800.0, 2600.0))
        burn(b["T"], b["p"], b["rho"], b["Y"], np.full(m, burn_time))
        b["rho"] = mixture_density(b["p"], b["T"], b["Y"], W)
        pick = rng.integers(0, m, size=n_burnt)
        parts.append({"T": b["T"][pick], "p": b["p"][pick], "rho": b["rho"][pick],
                      "Y": b["Y"][:, pick]})
    out = {k: np.concatenate([q[k] for q in parts], axis=-1) for k in ("T", "p", "rho", "Y")}
    perm = rng.permutation(n)  # interleave the kinds
    out = {k: np.ascontiguousarray(v[..., perm]) for k, v in out.items()}
    out["kind"] = np.concatenate([np.zeros(n_front, np.int8), np.ones(n_cold, np.int8),
                                  np.full(n_burnt, 2, np.int8)])[perm]
    return out
