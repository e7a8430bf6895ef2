"""Synthetic mechanisms (the 32- and 54-species workloads): generator
properties, and the oracle's rate laws on them -- fall-off (Lindemann, Troe,
SRI) Jacobian terms against finite differences, conservation invariants."""
import numpy as np
import pytest

RU = 8314.46261815324


@pytest.fixture(scope="module")
def syn32():
    from openhurricane_amd.chemistry import Mechanism

    return Mechanism.synthetic(32, 150, seed=1)


@pytest.fixture(scope="module")
def syn32_oracle(syn32):
    from oracle import Oracle

    return Oracle(syn32)


def test_generator_shape_and_determinism(syn32):
    from openhurricane_amd.chemistry import Mechanism

    t = syn32.tables()
    assert t["ns"] == 32 and t["nr"] == 150
    assert Mechanism.synthetic(32, 150, seed=1).hash == syn32.hash
    assert Mechanism.synthetic(32, 150, seed=2).hash != syn32.hash
    kinds = np.bincount(t["kind"], minlength=4)
    assert kinds[1] + kinds[2] + kinds[3] == 15  # 10 % third body
    assert kinds[2] + kinds[3] == 5              # a third of them fall-off
    assert set(t["fo_type"][t["kind"] >= 2]) >= {0, 2, 3}  # Lindemann, Troe, SRI
    assert (t["rev"] == 1).sum() == 30           # 20 % reversible
    m54 = Mechanism.synthetic(54, 270, seed=2).tables()
    assert m54["ns"] == 54 and m54["nr"] == 270


@pytest.mark.parametrize("ns,nr,seed", [(32, 150, 1), (54, 270, 2)])
def test_reactions_are_element_balanced(ns, nr, seed):
    from openhurricane_amd.chemistry import Mechanism

    t = Mechanism.synthetic(ns, nr, seed=seed).tables()
    E = t["elem"]
    for r in range(nr):
        f = sum(t["f_nu"][r][i] * E[t["f_idx"][r][i]] for i in range(t["nf"][r]))
        b = sum(t["b_nu"][r][i] * E[t["b_idx"][r][i]] for i in range(t["nb"][r]))
        np.testing.assert_allclose(f, b, err_msg=f"reaction {r}")


def test_thermo_is_smooth_and_continuous(syn32, syn32_oracle):
    """cp > 0 and rising over 300..5000 K; cp, h and s continuous at 1000 K."""
    Ts = np.linspace(300.0, 5000.0, 300)
    cp = np.array([syn32_oracle.species_thermo(T)["Cp"] for T in Ts])
    assert np.all(cp > 0)
    assert np.all(np.diff(cp, axis=0) > -1e-3 * cp[:-1])  # monotone up to fit ripple
    lo, hi = syn32_oracle.species_thermo(1000.0 - 1e-9), syn32_oracle.species_thermo(1000.0)
    for k in ("Cp", "H", "S"):
        np.testing.assert_allclose(lo[k], hi[k], rtol=1e-6, atol=1e-3, err_msg=k)


@pytest.mark.parametrize("flow", [0, 1])
def test_falloff_jacobian_matches_finite_differences(syn32, syn32_oracle, flow):
    W = syn32.tables()["W"]
    ns = 32
    n = ns + 1
    rng = np.random.default_rng(7 + flow)
    for _ in range(3):
        c = rng.uniform(0.05, 1.0, ns) * 10 ** rng.uniform(-6, -2, ns)
        y = np.concatenate([c, [rng.uniform(1000.0, 2400.0)]])
        pr = 2e5 if flow == 0 else float((c * W).sum())
        f, J = syn32_oracle.jacobian(y, pr, flow=flow, exact_dkp=True)
        Jfd = np.zeros((n, n))
        for j in range(n):
            h = 1e-6 * abs(y[j])
            yp, ym = y.copy(), y.copy()
            yp[j] += h
            ym[j] -= h
            Jfd[:, j] = (syn32_oracle.dydt(yp, pr, flow) - syn32_oracle.dydt(ym, pr, flow)) / (2 * h)
        row = np.abs(Jfd).max(axis=1, keepdims=True)
        live = row[:, 0] > 0
        err = np.abs(J - Jfd)[live] / row[live]
        assert err.max() < 1e-5, err.max()


def test_conservation_on_synthetic(syn32, syn32_oracle):
    t = syn32.tables()
    W, E = t["W"], t["elem"]
    rng = np.random.default_rng(3)
    for _ in range(10):
        T = rng.uniform(800, 2800)
        c = rng.uniform(0, 1, 32) * 10 ** rng.uniform(-7, -2, 32)
        w = syn32_oracle.omega(1e5, T, c)
        assert abs((w * W).sum()) <= 1e-12 * np.abs(w * W).sum()
        el = E.T @ w
        assert np.all(np.abs(el) <= 1e-12 * (np.abs(E.T) @ np.abs(w)) + 1e-300)
        y = np.concatenate([c, [T]])
        f = syn32_oracle.dydt(y, 1e5)
        th = syn32_oracle.species_thermo(T)
        lhs = (th["H"] * f[:-1]).sum() + (c * th["Cp"]).sum() * f[-1]
        assert abs(lhs) <= 1e-10 * np.abs(th["H"] * f[:-1]).sum()


def test_workload_sampler(syn32):
    from openhurricane_amd.sampler import synthetic_workload

    t = syn32.tables()
    calls = []

    def burn(T, p, rho, Y, dt):
        calls.append(len(T))
        T += 500.0  # stands in for an integrator

    st = synthetic_workload(syn32.species, t["W"], 1000, seed=3, burn=burn, pool=64)
    assert calls == [64]
    kind = st["kind"]
    assert (kind == 0).sum() == 200 and (kind == 2).sum() == 400
    front = st["T"][kind == 0]
    assert front.min() >= 1100 and front.max() <= 1600
    assert st["T"][kind == 2].min() >= 2300
    np.testing.assert_allclose(st["Y"].sum(axis=0), 1.0, rtol=1e-14)
