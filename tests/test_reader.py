"""CHEMKIN reader: auxiliary lines that carry several keyword groups.

The reference reads every keyword found on an auxiliary line
(chemkinFileRead.cpp:246-263, findLineKeywordType): LOW together with TROE
or SRI, LOW followed by third-body efficiencies, REV with DUP.  The same
mechanism written one keyword per line must give identical tables, and a
species whose name begins with a keyword (here LOWN2, an N2 renamed in both
files) is an efficiency, not that keyword."""
import numpy as np
import pytest

from openhurricane_amd.chemistry import DATA

HEAD = """ELEMENTS
O N H
END
SPECIES
O2 O H2 H2O OH H H2O2 HO2 LOWN2
END
REACTIONS
"""

ONE_LINE = HEAD + """ H + O2 (+M) <=> HO2 (+M)   4.65E12  0.44  0.0
LOW/5.75E19 -1.4 0.0/ TROE/0.5 1.0E-30 1.0E+30/ H2O/16.0/ LOWN2/1.3/
 H2O2 (+M) <=> OH + OH (+M)   2.0E12  0.9  48749.0
LOW/2.49E24 -2.3 48749.0/  SRI/0.8 300.0 900.0 1.2 0.1/ H2/3.7/ LOWN2/1.5/
 H + OH (+M) <=> H2O (+M)   1.0E14  0.0  0.0
HIGH/4.0E13 0.2 100.0/ TROE/0.6 100.0 2000.0 5000.0/
 H + OH + M = H2O + M	 8.62E+21  -2.0  0.0
LOWN2/0.4/ H2/2.5/ H2O/16.0/
 H2 + O2 = HO2 + H	  7.00E+13  0.0  56800
REV/1.0E13 0.1 200.0/ DUP
 H2 + O2 = HO2 + H	  1.00E+13  0.0  56000
DUPLICATE REV/2.0E12 0.0 100.0/
 O + H2 => OH + H	  5.06E+4  2.67  6290
FORD/H2 1.5/ FORD/O 0.8/
END
"""

MULTI_LINE = HEAD + """ H + O2 (+M) <=> HO2 (+M)   4.65E12  0.44  0.0
LOW/5.75E19 -1.4 0.0/
TROE/0.5 1.0E-30 1.0E+30/
H2O/16.0/ LOWN2/1.3/
 H2O2 (+M) <=> OH + OH (+M)   2.0E12  0.9  48749.0
LOW/2.49E24 -2.3 48749.0/
SRI/0.8 300.0 900.0 1.2 0.1/
H2/3.7/
LOWN2/1.5/
 H + OH (+M) <=> H2O (+M)   1.0E14  0.0  0.0
HIGH/4.0E13 0.2 100.0/
TROE/0.6 100.0 2000.0 5000.0/
 H + OH + M = H2O + M	 8.62E+21  -2.0  0.0
LOWN2/0.4/
H2/2.5/ H2O/16.0/
 H2 + O2 = HO2 + H	  7.00E+13  0.0  56800
REV/1.0E13 0.1 200.0/
DUP
 H2 + O2 = HO2 + H	  1.00E+13  0.0  56000
DUPLICATE
REV/2.0E12 0.0 100.0/
 O + H2 => OH + H	  5.06E+4  2.67  6290
FORD/H2 1.5/
FORD/O 0.8/
END
"""


def _files(tmp_path, text, tag):
    mech = tmp_path / f"mech_{tag}.che"
    mech.write_text(text)
    thermo = tmp_path / "thermo.dat"
    lines = (DATA / "h2air" / "thermo.dat").read_text().splitlines()
    out = []
    for ln in lines:  # rename N2 (columns 1-18 hold the name)
        if ln.startswith("N2 "):
            ln = "LOWN2" + ln[5:]
        out.append(ln)
    thermo.write_text("\n".join(out) + "\n")
    return mech, thermo


def _load(tmp_path, text, tag):
    from openhurricane_amd.chemistry import Mechanism

    mech, thermo = _files(tmp_path, text, tag)
    return Mechanism.load(mech, thermo, None)


def test_keywords_on_one_line_equal_one_per_line(tmp_path):
    a = _load(tmp_path, ONE_LINE, "one").tables()
    b = _load(tmp_path, MULTI_LINE, "multi").tables()
    assert a["nr"] == b["nr"] == 7
    for k, v in a.items():
        if isinstance(v, np.ndarray):
            np.testing.assert_array_equal(v, b[k], err_msg=k)
        else:
            assert v == b[k], k


def test_grouped_values_land_where_they_belong(tmp_path):
    m = _load(tmp_path, ONE_LINE, "one")
    t = m.tables()
    sp = m.species
    n2, h2o, h2 = sp.index("LOWN2"), sp.index("H2O"), sp.index("H2")
    # fall-off with TROE (3 parameters) and efficiencies from the LOW line
    assert t["kind"][0] == 2 and t["fo_type"][0] == 1
    np.testing.assert_array_equal(t["fo"][0][:3], [0.5, 1e-30, 1e30])
    assert t["eff"][0][h2o] == 16.0 and t["eff"][0][n2] == 1.3 and t["eff"][0][h2] == 1.0
    # fall-off with SRI (5 parameters)
    assert t["kind"][1] == 2 and t["fo_type"][1] == 3
    np.testing.assert_array_equal(t["fo"][1], [0.8, 300.0, 900.0, 1.2, 0.1])
    assert t["eff"][1][h2] == 3.7 and t["eff"][1][n2] == 1.5
    # chemically activated (HIGH) with a four-parameter TROE
    assert t["kind"][2] == 3 and t["fo_type"][2] == 2
    # a third-body line whose first species name begins with LOW
    assert t["kind"][3] == 1
    assert t["eff"][3][n2] == 0.4 and t["eff"][3][h2] == 2.5 and t["eff"][3][h2o] == 16.0
    # REV next to DUP / DUPLICATE, in either order
    assert t["rev"][4] == 2 and t["rev"][5] == 2
    assert t["arr_rev"][4][1] == 0.1 and t["arr_rev"][5][0] > 0
    # two FORD groups on one line
    o, h2i = sp.index("O"), sp.index("H2")
    f = {int(t["f_idx"][6][i]): t["f_ord"][6][i] for i in range(t["nf"][6])}
    assert f[h2i] == 1.5 and f[o] == 0.8


@pytest.mark.parametrize("line", ["LOW/1 2/ TROE 0.5 1 2", "H2/2.5", "BOGUS", "PLOG/1.0 1e13 0 0/"])
def test_malformed_or_unsupported_groups_are_errors(tmp_path, line):
    from openhurricane_amd._lib import OhcError

    text = HEAD + " H + O2 (+M) <=> HO2 (+M)   4.65E12  0.44  0.0\n" + line + "\nEND\n"
    with pytest.raises(OhcError):
        _load(tmp_path, text, "bad")


# ---- PLOG and CHEB (chemkinFileRead.cpp:606-755, units :1595-1627) -------
RATES = HEAD + """ H + O2 <=> O + OH   1.0 0.0 0.0
PLOG / 10.0  3.0E13 0.1 1000.0 /
PLOG / 0.1   1.0E13 0.0 500.0 /
PLOG / 1.0   2.0E13 0.0 700.0 /  PLOG / 1.0 5.0E12 0.0 800.0 /
 H2O2 (+M) <=> OH + OH (+M)   1.0 0.0 0.0
TCHEB / 300.0 2200.0 /
PCHEB / 0.01 98.7 /
CHEB / 3 2 1.0 2.0 /
CHEB / 3.0 4.0 5.0 6.0 /
 H + OH (+M) <=> H2O (+M)   1.0 0.0 0.0
CHEB / 2 2 0.5 0.1 0.2 0.3 /  H2O / 5.0 /
 O + H2 <=> OH + H   5.06E+4  2.67  6290
END
"""
PATM = 101325.0
R_CAL = 8.314462618 / 4.184  # cal/(mol K)


def test_plog_points_sorted_and_in_si_units(tmp_path):
    m = _load(tmp_path, RATES, "rates")
    t = m.tables()
    assert list(t["kind"]) == [4, 5, 5, 0]
    d = m.rate_data(0)
    assert d["law"] == "PLOG"
    # ascending pressure; equal pressures keep the order they were read in
    exp = [[0.1, 1.0e13, 0.0, 500.0], [1.0, 2.0e13, 0.0, 700.0],
           [1.0, 5.0e12, 0.0, 800.0], [10.0, 3.0e13, 0.1, 1000.0]]
    for got, (p, A, b, E) in zip(d["points"], exp):
        # bimolecular: A from cm^3/(mol s) to m^3/(kmol s) is x 1e-3
        np.testing.assert_allclose(got, [p * PATM, A * 1e-3, b, E / R_CAL], rtol=1e-9)
    assert m.rate_data(3) == {}


def test_cheb_grid_limits_and_defaults(tmp_path):
    m = _load(tmp_path, RATES, "rates")
    d = m.rate_data(1)
    assert d["law"] == "CHEB" and (d["N"], d["M"]) == (3, 2)
    assert d["T"] == (300.0, 2200.0)
    np.testing.assert_allclose(d["P"], (0.01 * PATM, 98.7 * PATM), rtol=1e-12)
    np.testing.assert_array_equal(d["coef"], [[1, 2], [3, 4], [5, 6]])
    d2 = m.rate_data(2)
    assert (d2["N"], d2["M"]) == (2, 2) and d2["T"] == (300.0, 2500.0)
    np.testing.assert_allclose(d2["P"], (0.001 * PATM, 100.0 * PATM), rtol=1e-12)
    # efficiencies of a (+M) CHEB reaction are kept for M
    t = m.tables()
    assert t["eff"][2][m.species.index("H2O")] == 5.0


def test_plog_cheb_round_trip_through_tables(tmp_path):
    from openhurricane_amd.chemistry import Mechanism

    m = _load(tmp_path, RATES, "rates")
    t = m.tables()
    m2 = Mechanism.from_tables(t)
    t2 = m2.tables()
    np.testing.assert_array_equal(t2["aux_off"], t["aux_off"])
    np.testing.assert_array_equal(t2["aux"], t["aux"])
    assert m2.hash == m.hash
    # mechanisms without PLOG/CHEB keep their hash: the tables carry no aux data
    h2 = Mechanism.h2air()
    assert h2.tables()["aux_off"][-1] == 0


def test_plog_cheb_mechanism_is_refused_by_the_gpu_path(tmp_path):
    """The reference's GPU table builder stops on these laws
    (reversibleReactions.cpp:186, "Unsupported reaction rate type"); the
    context refuses the mechanism with OHC_ERR_UNSUPPORTED before it touches
    a device, and the oracle (which restates the GPU path) refuses it too."""
    from oracle import Oracle

    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource

    m = _load(tmp_path, RATES, "rates")
    with pytest.raises(_lib.OhcError) as e:
        ChemistrySource(m)
    assert e.value.code == _lib.OHC_ERR_UNSUPPORTED and "PLOG" in str(e.value)
    with pytest.raises(ValueError):
        Oracle(m)


@pytest.mark.parametrize("rxn,aux", [
    (" H + O2 <=> O + OH   1.0 0.0 0.0", "PLOG / 1.0 2.0E13 0.0 /"),           # 3 numbers
    (" H + O2 <=> O + OH   1.0 0.0 0.0", "PLOG / 1.0 2E13 0 0 / REV / 1 0 0 /"),
    (" H + O2 (+M) <=> HO2 (+M)   1.0 0.0 0.0", "PLOG / 1.0 2E13 0 0 /"),     # (+M)
    (" H + O2 <=> O + OH   1.0 0.0 0.0", "CHEB / 2 2 1 2 3 4 /"),             # no (+M)
    (" H + O2 (+M) <=> HO2 (+M)   1.0 0.0 0.0", "CHEB / 2 2 1 2 3 /"),        # count
    (" H + O2 (+M) <=> HO2 (+M)   1.0 0.0 0.0", "TCHEB / 300 2500 /"),        # no grid
    (" H + O2 (+M) <=> HO2 (+M)   1.0 0.0 0.0", "CHEB / 1 1 1 / LOW / 1 0 0 /"),
    (" H + O2 (+M) <=> HO2 (+M)   1.0 0.0 0.0", "CHEB / 1 1 1 / TCHEB / 900 300 /"),
    (" H + O2 <=> O + OH   1.0 0.0 0.0", "PLOG / 1 2E13 0 0 / CHEB / 1 1 1 /"),
])
def test_malformed_plog_cheb_are_errors(tmp_path, rxn, aux):
    from openhurricane_amd._lib import OhcError

    with pytest.raises(OhcError):
        _load(tmp_path, HEAD + rxn + "\n" + aux + "\nEND\n", "bad")


def test_text_mechanism_rates_match_closed_forms(tmp_path):
    """LOW/TROE/SRI/HIGH/REV/FORD read from CHEMKIN text, evaluated by the
    oracle, against the textbook rate forms computed here from the numbers
    as written in the text: unit factors of chemkinFileRead.cpp:1654-1699
    (A x 1e-3 per order above one, x 1e-3 more for k0 of a fall-off and for a
    +M third body, x 1e+3 for the HIGH k_inf of a chemically activated
    reaction, :1448-1482), E in cal/mol; Troe and SRI in their published
    forms (Gilbert, Luther and Troe 1983; Kee et al., CHEMKIN-II manual)."""
    from oracle import Oracle

    m = _load(tmp_path, ONE_LINE, "one")
    sp = m.species
    orc = Oracle(m)
    rng = np.random.default_rng(5)
    c = rng.uniform(0.5e-3, 5e-3, len(sp))
    ix = {s: sp.index(s) for s in sp}
    for T in (700.0, 1350.0, 2600.0):
        r = orc.rates(1e5, T, c)
        arr = lambda A, b, E: A * T ** b * np.exp(-E / (R_CAL * T))  # noqa: E731
        ctot = c.sum()

        def M(**eff):
            return ctot + sum((v - 1.0) * c[ix[k]] for k, v in eff.items())

        def troe(Pr, Fc):
            lf = np.log10(Fc)
            cc, nn = -0.4 - 0.67 * lf, 0.75 - 1.27 * lf
            f1 = (np.log10(Pr) + cc) / (nn - 0.14 * (np.log10(Pr) + cc))
            return 10.0 ** (lf / (1.0 + f1 * f1))

        exp = np.zeros(7)
        # 0: Lindemann-Troe fall-off, TROE/0.5 1e-30 1e30/
        kinf, k0 = arr(4.65e12 * 1e-3, 0.44, 0.0), arr(5.75e19 * 1e-6, -1.4, 0.0)
        Pr = k0 * M(H2O=16.0, LOWN2=1.3) / kinf
        Fc = 0.5 * np.exp(-T / 1e-30) + 0.5 * np.exp(-T / 1e30)
        exp[0] = kinf * Pr / (1 + Pr) * troe(Pr, Fc)
        # 1: unimolecular fall-off with SRI/0.8 300 900 1.2 0.1/
        kinf, k0 = arr(2.0e12, 0.9, 48749.0), arr(2.49e24 * 1e-3, -2.3, 48749.0)
        Pr = k0 * M(H2=3.7, LOWN2=1.5) / kinf
        X = 1.0 / (1.0 + np.log10(Pr) ** 2)
        F = 1.2 * (0.8 * np.exp(-300.0 / T) + np.exp(-T / 900.0)) ** X * T ** 0.1
        exp[1] = kinf * Pr / (1 + Pr) * F
        # 2: chemically activated, main line k0, HIGH k_inf, TROE 4 parameters
        k0, kinf = arr(1.0e14 * 1e-3, 0.0, 0.0), arr(4.0e13, 0.2, 100.0)
        Pr = k0 * M() / kinf
        Fc = 0.4 * np.exp(-T / 100.0) + 0.6 * np.exp(-T / 2000.0) + np.exp(-5000.0 / T)
        exp[2] = k0 / (1 + Pr) * troe(Pr, Fc)
        # 3: +M third body with efficiencies
        exp[3] = arr(8.62e21 * 1e-6, -2.0, 0.0) * M(LOWN2=0.4, H2=2.5, H2O=16.0)
        # 4, 5: explicit reverse parameters (REV next to DUP)
        exp[4], exp[5] = arr(7.0e13 * 1e-3, 0.0, 56800.0), arr(1.0e13 * 1e-3, 0.0, 56000.0)
        np.testing.assert_allclose(r["kr"][4], arr(1.0e13 * 1e-3, 0.1, 200.0), rtol=1e-9)
        np.testing.assert_allclose(r["kr"][5], arr(2.0e12 * 1e-3, 0.0, 100.0), rtol=1e-9)
        # 6: FORD H2 1.5, O 0.8: order 2.3 sets the A factor and the exponents
        exp[6] = arr(5.06e4 * 1e-3 ** 1.3, 2.67, 6290.0)
        np.testing.assert_allclose(r["qf"][6], exp[6] * c[ix["H2"]] ** 1.5 * c[ix["O"]] ** 0.8,
                                   rtol=1e-9)
        assert r["kr"][6] == 0.0
        np.testing.assert_allclose(r["kf"], exp, rtol=1e-9, err_msg=f"T={T}")
