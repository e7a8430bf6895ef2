"""Pins of the oracle's MTS explicit integrator (the reference GPU path's
MTSEuler solver: solveChemODEsMTS, CUDAChemistrySourceODEs.cu:684-778, with
timescale / MTS_whichGroup / MTS_GroupDT :199-309 and
reactionTable::dWdciAllSpecies, reactionTableCUDA.cu:331-385).

* Time scales: 1 / |d omega_k / d c_k| at fixed T -- the diagonal of the
  (finite-difference pinned) analytic Jacobian's species block, and central
  differences of the net rates where rounding allows -- on a mechanism
  without third-body reactions, whose efficiency terms the reference's
  diagonal leaves out.
* The isomerisation A -> B at rate k is solved in closed form: the fast group
  {A} relaxes by the factor 1 - 0.8 per explicit step until the increment is
  within tolerance, the empty intermediate groups are skipped, and the last
  group advances B alone to the end of the step with A frozen.
* A cell whose time scales all exceed the flow step forms one group and is
  one forward-Euler step: its error against a tight Radau solution falls as
  the square of the step.
"""
import numpy as np
import pytest


def _iso_mech(k):
    from openhurricane_amd.chemistry import Mechanism

    P = 6
    nasa = np.tile([3.5, 0, 0, 0, 0, -1000.0, 5.0], (2, 1))
    z = lambda *s, dt=float: np.zeros(s, dtype=dt)
    f_idx, b_idx = z(1, P, dt=np.int32), z(1, P, dt=np.int32)
    b_idx[0, 0] = 1
    nu = z(1, P)
    nu[0, 0] = 1.0
    tab = {"ns": 2, "nr": 1, "ne": 1, "W": np.array([28.0, 28.0]),
           "Tlow": np.full(2, 200.0), "Thigh": np.full(2, 6000.0), "Tcommon": np.full(2, 1000.0),
           "nasa_lo": nasa, "nasa_hi": nasa, "elem": np.array([[2.0], [2.0]]),
           "nf": np.array([1]), "nb": np.array([1]), "f_idx": f_idx, "b_idx": b_idx,
           "f_nu": nu, "b_nu": nu, "f_ord": nu, "b_ord": nu, "rev": np.array([0]),
           "kind": np.array([0]), "arr": np.array([[k, 0.0, 0.0]]), "arr_lo": z(1, 3),
           "arr_rev": z(1, 3), "fo_type": np.array([0]), "fo": z(1, 5), "eff": z(1, 2)}
    return Mechanism.from_tables(tab)


def test_timescales_are_inverse_diagonal_rates():
    from oracle import Oracle

    from openhurricane_amd.chemistry import Mechanism
    from openhurricane_amd.sampler import synthetic_workload

    mech = Mechanism.synthetic(32, 150, seed=7, frac_rev=0.2, frac_tb=0.0)
    orc = Oracle(mech)
    W = mech.tables()["W"]
    st = synthetic_workload(mech.species, W, 6, seed=3, front=0.5)
    for i in range(6):
        T, p = st["T"][i], st["p"][i]
        c = st["rho"][i] * st["Y"][:, i] / W
        tci = orc.mts_timescales(T, c)
        # the species block of the analytic Jacobian (finite-difference pinned
        # in test_oracle_pins) holds d omega_k / d c_k at fixed T
        _, J = orc.jacobian(np.concatenate([c, [T]]), p)
        want = np.minimum(200.0, 1.0 / (np.abs(np.diag(J)[:32]) + 1e-20))
        np.testing.assert_allclose(tci, want, rtol=1e-12)
        # and, where the net rate is small enough for a clean difference,
        # directly by central differences of the net rates
        for kk in range(32):
            h = 1e-3 * c[kk]
            if not h > 0:
                continue
            cp_, cm = c.copy(), c.copy()
            cp_[kk] += h
            cm[kk] -= h
            wp, wm, w0 = orc.omega(p, T, cp_)[kk], orc.omega(p, T, cm)[kk], orc.omega(p, T, c)[kk]
            d = (wp - wm) / (2 * h)
            if abs(w0) * 1e-15 / h > 1e-6 * abs(d):  # rounding would dominate
                continue
            assert abs(min(200.0, 1.0 / (abs(d) + 1e-20)) - tci[kk]) <= 1e-6 * tci[kk], kk


@pytest.mark.parametrize("k,c0", [(3e9, 1e-2), (5.5e8, 3e-3)])
def test_isomerisation_closed_form(k, c0):
    from oracle import Oracle

    from openhurricane_amd import _lib

    mech = _iso_mech(k)
    orc = Oracle(mech)
    atol, rtol, tn, cB0 = 1e-11, 1e-4, 1e-6, 2e-3
    o = _lib.default_opts(atol=atol, rtol=rtol, integrator=4)
    tci = orc.mts_timescales(1000.0, np.array([c0, cB0]))
    np.testing.assert_allclose(tci, [1.0 / k, 200.0], rtol=1e-14)
    y, groups, steps = orc.mts_cell(o, [c0, cB0, 1000.0], 101325.0, tn)
    assert groups == int(np.floor(np.log10(tn * k))) + 1
    # fast group: step j is taken while the previous step's increment
    # 0.8 c_(j-1) exceeded its tolerance atol + rtol c_(j-1)
    m, c = 0, c0
    while True:
        m += 1
        inc = 0.8 * c
        c = c - inc
        if inc / (atol + rtol * max(c, c + inc)) < 1.0:
            break
    cA = c0 * 0.2 ** m
    cB = cB0 + c0 * (1 - 0.2 ** m) + (tn - m * 0.8 / k) * k * cA
    assert steps == m + 1
    np.testing.assert_allclose(y[:2], [cA, cB], rtol=1e-12)
    assert y[2] == 1000.0


def test_single_group_is_second_order(h2mech, h2oracle):
    from scipy.integrate import solve_ivp

    from openhurricane_amd import _lib
    from openhurricane_amd.sampler import h2air_states

    W = h2mech.tables()["W"]
    st = h2air_states(h2mech.species, W, 4, seed=5, T_range=(1000, 1200))
    o = _lib.default_opts(atol=1e-11, rtol=1e-4, integrator=4)
    for i in range(4):
        y0 = np.concatenate([st["rho"][i] * st["Y"][:, i] / W, [st["T"][i]]])
        p = st["p"][i]
        tmin = h2oracle.mts_timescales(y0[-1], y0[:-1]).min()
        errs = []
        for tn in (0.2 * tmin, 0.1 * tmin):
            y, groups, steps = h2oracle.mts_cell(o, y0, p, tn)
            assert groups == 1 and steps == 1
            sol = solve_ivp(lambda t, yy: h2oracle.dydt(yy, p), (0, tn), y0, method="Radau",
                            rtol=1e-13, atol=1e-24)
            ref = sol.y[:, -1]
            errs.append(np.max(np.abs(y - ref) / np.maximum(np.abs(ref), 1e-6)))
        assert 3.0 < errs[0] / errs[1] < 5.0, errs
