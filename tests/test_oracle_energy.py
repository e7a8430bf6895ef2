"""Pins of the oracle's internal-energy forms (thermoList::TEa_p / TEa_rho,
thermoList.inl:343-370; cv_p and ea_p, thermoList.inl:145-156, :289-298):

* ea = ha - p/rho with rho = p W / (Ru T) computed here from the species
  molecular weights (the perfect-gas equation of state, perfectGas.cpp:36-47);
* cv = d(ea)/dT by central finite differences, and cv = cp - Ru/W;
* T(e) inverts e(T) to the Newton stopping tolerance, with the limiter's
  clamps and flags outside [T_low, T_high]."""
import numpy as np

RU = 8314.46261815324


def _states(h2mech, n, seed, T_range=(300.0, 3500.0)):
    from openhurricane_amd.sampler import h2air_states

    return h2air_states(h2mech.species, h2mech.tables()["W"], n, seed=seed, T_range=T_range)


def test_ea_and_cv_closed_forms(h2mech, h2oracle):
    W = h2mech.tables()["W"]
    st = _states(h2mech, 64, 4)
    th = h2oracle.mix_thermo(st["T"], st["p"], st["Y"])
    Wm = 1.0 / (st["Y"] / W[:, None]).sum(0)
    rho = st["p"] * Wm / (RU * st["T"])
    np.testing.assert_allclose(th["ea"], th["ha"] - st["p"] / rho, rtol=1e-12, atol=1e-6)
    np.testing.assert_allclose(th["cv"], th["cp"] - RU / Wm, rtol=1e-13)
    # cv = d(ea)/dT (central differences, away from the NASA-7 switch at 1000 K)
    T = st["T"]
    mask = np.abs(T - 1000.0) > 5.0
    h = 1e-3 * T
    ep = h2oracle.mix_thermo(T + h, st["p"], st["Y"])["ea"]
    em = h2oracle.mix_thermo(T - h, st["p"], st["Y"])["ea"]
    np.testing.assert_allclose(((ep - em) / (2 * h))[mask], th["cv"][mask], rtol=1e-6)
    # the mixture forms equal the mass-weighted species values
    sp = np.array([h2oracle.species_thermo(t)["H"] for t in T]).T / W[:, None]
    np.testing.assert_allclose(th["ha"], (sp * st["Y"]).sum(0), rtol=1e-12, atol=1e-6)


def test_T_from_e_round_trip_and_clamps(h2mech, h2oracle, c1opts):
    st = _states(h2mech, 256, 5)
    th = h2oracle.mix_thermo(st["T"], st["p"], st["Y"])
    T, fl = h2oracle.T_from_e(c1opts, th["ea"], st["p"], st["Y"], np.full(256, 1500.0))
    np.testing.assert_allclose(T, st["T"], rtol=2e-4)
    assert np.all(fl == 1)
    back = h2oracle.mix_thermo(T, st["p"], st["Y"])["ea"]
    assert np.all(np.abs(back - th["ea"]) <= 1e-4 * 1500.0 * th["cv"] * 1e-2)
    # e(T) is not h(T): inverting e with the enthalpy Newton misses the state
    Th, _ = h2oracle.T_from_h(c1opts, th["ea"], st["p"], st["Y"], np.full(256, 1500.0))
    assert np.max(np.abs(Th - st["T"]) / st["T"]) > 1e-2
    sub = {k: np.ascontiguousarray(v[..., :4]) for k, v in st.items()}
    lo = h2oracle.mix_thermo(np.full(4, 0.4 * c1opts.T_low), sub["p"], sub["Y"])["ea"]
    T, fl = h2oracle.T_from_e(c1opts, lo, sub["p"], sub["Y"], np.full(4, 800.0))
    assert np.all(T == c1opts.T_low) and np.all(fl == 0)
    hi = h2oracle.mix_thermo(np.full(4, 1.2 * c1opts.T_high), sub["p"], sub["Y"])["ea"]
    T, fl = h2oracle.T_from_e(c1opts, hi, sub["p"], sub["Y"], np.full(4, 800.0))
    assert np.all(T == c1opts.T_high) and np.all(fl == 2)
