"""Seeded synthetic cell states for the benchmark and parity workloads.

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
