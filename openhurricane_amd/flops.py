"""Algorithmic work model of the plain stiff integrator (per cell-advance).

Counts FP64 floating-point operations of the reference algorithm from the
mechanism structure, as
  work = N_f F_rhs + N_J (F_J + F_rhs + 2 n^2) + N_LU (2 n^3 / 3) + N_BS (2 n^2)
with N_J = N_LU = N_BS = Newton iterations and N_f = extra RHS evaluations
(the first-step error estimate); transcendentals (exp, log, pow) are counted
separately and not included in the flop total.
"""
from __future__ import annotations


def integrator_flops(tab: dict) -> dict:
    ns, nr = tab["ns"], tab["nr"]
    n = ns + 1
    F_rhs = 14 * ns + 4 * ns
    F_J = 4 * ns * ns + 8 * ns
    T_rhs = ns
    for r in range(nr):
        m = int(tab["nf"][r] + tab["nb"][r])
        tb = tab["kind"][r] != 0
        rev = tab["rev"][r] != 0
        F_rhs += 2 * m + 6 + (2 * ns if tb else 0) + (2 * m + 2 if rev else 0)
        F_J += m * m + 2 * m * m + (2 * ns * m if tb else 0)
        T_rhs += 2 + (1 if rev else 0)
    per_newton = F_J + F_rhs + 2 * n * n + 2 * n ** 3 / 3.0 + 2 * n * n
    return {"F_rhs": F_rhs, "F_J": F_J, "F_newton": per_newton, "T_rhs": T_rhs, "n": n}


def work_flops(tab: dict, stats: dict) -> float:
    f = integrator_flops(tab)
    return stats["newton_iters"] * f["F_newton"] + stats["rhs_evals"] * f["F_rhs"]
