"""Pins of the oracle's species thermo (JANAF NASA-7) by closed forms,
thermodynamic identities, finite differences and external table values."""
import numpy as np
import pytest

RU = 8314.46261815324  # J/kmol/K


def nasa(raw_lo, raw_hi, Tc, T):
    a = raw_lo if T < Tc else raw_hi
    cp = a[0] + a[1] * T + a[2] * T**2 + a[3] * T**3 + a[4] * T**4
    h = a[0] + a[1] * T / 2 + a[2] * T**2 / 3 + a[3] * T**3 / 4 + a[4] * T**4 / 5 + a[5] / T
    s = a[0] * np.log(T) + a[1] * T + a[2] * T**2 / 2 + a[3] * T**3 / 3 + a[4] * T**4 / 4 + a[6]
    return cp, h, s  # cp/R, h/(RT), s/R


@pytest.mark.parametrize("T", [300.0, 650.0, 999.0, 1001.0, 1800.0, 3500.0, 5000.0])
def test_textbook_nasa_polynomials(h2mech, h2oracle, T):
    t = h2mech.tables()
    th = h2oracle.species_thermo(T)
    for k in range(t["ns"]):
        cp, h, s = nasa(t["nasa_lo"][k], t["nasa_hi"][k], t["Tcommon"][k], T)
        assert th["Cp"][k] == pytest.approx(cp * RU, rel=1e-12)
        assert th["H"][k] == pytest.approx(h * RU * T, rel=1e-11, abs=1e-6 * RU * T)
        assert th["S"][k] == pytest.approx(s * RU, rel=1e-11)
        # G = H - T S (JANAF g0 is evaluated from its own closed form)
        assert th["G"][k] == pytest.approx(th["H"][k] - T * th["S"][k],
                                           rel=1e-9, abs=1e-9 * RU * T)


@pytest.mark.parametrize("T", [400.0, 800.0, 1500.0, 2500.0, 4000.0])
def test_derivative_identities(h2oracle, T):
    d = 1e-3 * T
    a, b, c = h2oracle.species_thermo(T - d), h2oracle.species_thermo(T), h2oracle.species_thermo(T + d)
    # cp = dh/dT, dCp/dT, dG/dT = -S
    np.testing.assert_allclose((c["H"] - a["H"]) / (2 * d), b["Cp"], rtol=1e-6)
    np.testing.assert_allclose((c["Cp"] - a["Cp"]) / (2 * d), b["dCp"], rtol=1e-5, atol=1e-6 * RU)
    np.testing.assert_allclose(b["dG"], -b["S"], rtol=1e-10)
    np.testing.assert_allclose((c["G"] - a["G"]) / (2 * d), b["dG"], rtol=1e-6)


def test_out_of_range_rules(h2mech, h2oracle):
    """cp is clamped to [Tlow, Thigh]; h is extrapolated linearly with that cp."""
    lo = h2oracle.species_thermo(300.0)
    th = h2oracle.species_thermo(250.0)
    np.testing.assert_allclose(th["Cp"], lo["Cp"], rtol=1e-14)
    np.testing.assert_allclose(th["H"], lo["H"] + lo["Cp"] * (250.0 - 300.0), rtol=1e-12)
    hi = h2oracle.species_thermo(5000.0)
    th = h2oracle.species_thermo(5400.0)
    np.testing.assert_allclose(th["H"], hi["H"] + hi["Cp"] * 400.0, rtol=1e-12)


def test_external_table_values(h2mech, h2oracle):
    """Molar cp against JANAF-table values (J/mol/K) -- pins units and scaling."""
    sp = h2mech.species
    ref = {("H2O", 1000.0): 41.27, ("N2", 1000.0): 32.70, ("O2", 300.0): 29.39,
           ("H2", 2000.0): 34.28, ("OH", 1500.0): 32.95}
    for (name, T), cp in ref.items():
        th = h2oracle.species_thermo(T)
        assert th["Cp"][sp.index(name)] / 1e3 == pytest.approx(cp, rel=0.01), name
