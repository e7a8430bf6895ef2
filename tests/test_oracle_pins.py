"""Pins of the oracle's Jacobian, LU solve, integrator, T(h) inversion and
transport against independent checks: finite differences, numpy's LAPACK
solve, scipy's Radau integrator, round trips, and the textbook
Chapman-Enskog forms evaluated here from the raw transport file."""
import numpy as np
import pytest

RU = 8314.46261815324
PATM = 101325.0


def _state(rng, T_range=(1000.0, 2500.0)):
    T = rng.uniform(*T_range)
    c = rng.uniform(0, 1, 9) * 10 ** rng.uniform(-5, -1.5, 9)
    return np.concatenate([c, [T]])


@pytest.mark.parametrize("flow", [0, 1])
def test_jacobian_matches_finite_differences(h2mech, h2oracle, flow):
    """Exact-derivative variant against central differences of dy/dt; the
    reference's dKp/dT convention changes only the temperature column."""
    W = h2mech.tables()["W"]
    rng = np.random.default_rng(3 + flow)
    n = 10
    for _ in range(6):
        y = _state(rng)
        pr = 1e5 if flow == 0 else float((y[:-1] * W).sum())
        f, J = h2oracle.jacobian(y, pr, flow=flow, exact_dkp=True)
        Jfd = np.zeros((n, n))
        for j in range(n):
            h = 1e-6 * max(abs(y[j]), 1e-8) if j < n - 1 else 1e-4 * y[-1]
            yp, ym = y.copy(), y.copy()
            yp[j] += h
            ym[j] -= h
            Jfd[:, j] = (h2oracle.dydt(yp, pr, flow) - h2oracle.dydt(ym, pr, flow)) / (2 * h)
        row = np.abs(Jfd).max(axis=1, keepdims=True)
        live = row[:, 0] > 0  # the inert N2 row is identically zero
        assert np.all(J[~live] == 0.0)
        err = np.abs(J - Jfd)[live] / row[live]
        assert err.max() < 1e-6, err.max()
        np.testing.assert_allclose(f, h2oracle.dydt(y, pr, flow), rtol=1e-14)
        _, Jref = h2oracle.jacobian(y, pr, flow=flow, exact_dkp=False)
        np.testing.assert_array_equal(Jref[:, :-1], J[:, :-1])


@pytest.mark.parametrize("n", [10, 33, 55])
def test_lu_solve_matches_lapack(n):
    from oracle import Oracle

    rng = np.random.default_rng(n)
    for trial in range(5):
        A = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n) * (trial % 2)
        b = rng.standard_normal(n)
        np.testing.assert_allclose(Oracle.lu_solve(A, b), np.linalg.solve(A, b),
                                   rtol=1e-9, atol=1e-11)
    # pivoting is required: zero leading diagonal and a permutation matrix
    P = np.eye(n)[rng.permutation(n)]
    P[0, 0] = 0.0
    if not P[:, 0].any():
        P[1, 0] = 1.0
    A = P + 1e-3 * rng.standard_normal((n, n))
    A[0, 0] = 0.0
    b = rng.standard_normal(n)
    np.testing.assert_allclose(Oracle.lu_solve(A, b), np.linalg.solve(A, b), rtol=1e-9,
                               atol=1e-10)


def test_integrator_converges_to_radau(h2mech, h2oracle):
    """ImplicitEuler over dt = 1e-6 s approaches a tight Radau solution of the
    same ODE as the step tolerance tightens (backward Euler is first order:
    the error falls about 10x per 100x in tolerance)."""
    from scipy.integrate import solve_ivp

    from openhurricane_amd import _lib
    from openhurricane_amd.sampler import h2air_states

    W = h2mech.tables()["W"]
    n, dt = 6, 1e-6
    st = h2air_states(h2mech.species, W, n, seed=11, T_range=(1600, 2500))
    refT, refY = np.zeros(n), np.zeros((9, n))
    for i in range(n):
        y0 = np.concatenate([st["rho"][i] * st["Y"][:, i] / W, [st["T"][i]]])
        p = st["p"][i]
        sol = solve_ivp(lambda t, y: h2oracle.dydt(y, p), (0, dt), y0, method="Radau",
                        rtol=1e-12, atol=1e-22,
                        jac=lambda t, y: h2oracle.jacobian(y, p, exact_dkp=True)[1])
        y = sol.y[:, -1]
        refT[i] = y[-1]
        refY[:, i] = y[:-1] * W / (y[:-1] * W).sum()
    errs = []
    for rtol in (1e-4, 1e-6, 1e-8):
        o = _lib.default_opts(atol=1e-16, rtol=rtol, atol_inc=1e-16, rtol_inc=rtol)
        T, Y, _, _ = h2oracle.solve_update(o, st["T"], st["p"], st["rho"], st["Y"],
                                           np.full(n, dt), 1.0, subDt=np.full(n, 1e-13))
        errs.append(max(np.max(np.abs(T - refT) / refT),
                        np.max(np.abs(Y - refY) / np.maximum(refY, 1e-3))))
    assert errs[0] > 4 * errs[1] > 16 * errs[2], errs
    assert errs[2] < 3e-4, errs


def test_T_from_h_round_trip_and_clamps(h2mech, h2oracle, c1opts):
    from openhurricane_amd.sampler import h2air_states

    W = h2mech.tables()["W"]
    st = h2air_states(h2mech.species, W, 256, seed=2, T_range=(300, 3500))
    props = h2oracle.mix_props(st["T"], st["p"], st["Y"])
    T, fl = h2oracle.T_from_h(c1opts, props["ha"], st["p"], st["Y"], np.full(256, 1500.0))
    np.testing.assert_allclose(T, st["T"], rtol=2e-4)
    assert np.all(fl == 1)
    # h(T(h)) = h to the Newton stopping tolerance (|dT| <= 1e-4 T0, quadratic convergence)
    back = h2oracle.mix_props(T, st["p"], st["Y"])["ha"]
    cp = props["cp"]
    assert np.all(np.abs(back - props["ha"]) <= 1e-4 * 1500.0 * cp * 1e-2)
    # enthalpies outside [T_low, T_high] return the clamp and its flag
    st = {k: np.ascontiguousarray(v[..., :4]) for k, v in st.items()}
    lo = h2oracle.mix_props(np.full(4, 0.4 * c1opts.T_low), st["p"], st["Y"])["ha"]
    T, fl = h2oracle.T_from_h(c1opts, lo, st["p"], st["Y"], np.full(4, 800.0))
    assert np.all(T == c1opts.T_low) and np.all(fl == 0)
    hi = h2oracle.mix_props(np.full(4, 1.2 * c1opts.T_high), st["p"], st["Y"])["ha"]
    T, fl = h2oracle.T_from_h(c1opts, hi, st["p"], st["Y"], np.full(4, 800.0))
    assert np.all(T == c1opts.T_high) and np.all(fl == 2)


def _raw_transport():
    from openhurricane_amd.chemistry import DATA

    out = {}
    for line in (DATA / "h2air" / "transport.dat").read_text().splitlines():
        tok = line.split("!")[0].split()
        if len(tok) >= 7:
            out[tok[0]] = (float(tok[2]), float(tok[3]))  # eps/k [K], sigma [Angstrom]
    return out


def _omega22(Ts):
    return 1.147 * Ts ** -0.145 + (Ts + 0.5) ** -2


def test_transport_closed_forms(h2mech, h2oracle):
    """Pure species: mu = 2.6693e-6 sqrt(W T)/(sigma^2 Omega22) Pa s and
    kappa = mu R (cp/R + 1.25) (modified Eucken); a binary mixture's
    mixture-averaged D equals the binary D_12 = 1.858e-7 T^1.5
    sqrt(1/W1 + 1/W2) (p_atm/p)/(sigma_12^2 Omega_D) m^2/s."""
    sp = h2mech.species
    t = h2mech.tables()
    W = t["W"]
    raw = _raw_transport()
    T = np.array([300.0, 800.0, 1500.0, 2500.0])
    p = np.full(4, 2.0 * PATM)
    for k, name in enumerate(sp):
        Y = np.zeros((9, 4))
        Y[k] = 1.0
        out = h2oracle.mix_props(T, p, Y)
        eps, sig = raw[name]
        mu = 2.6693e-6 * np.sqrt(W[k] * T) / (sig ** 2 * _omega22(T / eps))
        np.testing.assert_allclose(out["mu"], mu, rtol=1e-13, err_msg=name)
        cpk = np.array([h2oracle.species_thermo(x)["Cp"][k] for x in T]) / W[k]
        Rk = RU / W[k]
        np.testing.assert_allclose(out["kappa"], mu * Rk * (cpk / Rk + 1.25), rtol=1e-13)
        np.testing.assert_allclose(out["cp"], cpk, rtol=1e-14)
    i, j = sp.index("H2"), sp.index("N2")
    Y = np.zeros((9, 4))
    Y[i], Y[j] = 0.3, 0.7
    out = h2oracle.mix_props(T, p, Y)
    (ei, si), (ej, sj) = raw["H2"], raw["N2"]
    OmD = (T / np.sqrt(ei * ej)) ** -0.145 + (T / np.sqrt(ei * ej) + 0.5) ** -2
    D12 = 1.858e-7 * T ** 1.5 * np.sqrt(1 / W[i] + 1 / W[j]) * PATM / (
        p * (0.5 * (si + sj)) ** 2 * OmD)
    np.testing.assert_allclose(out["Dk"][i], D12, rtol=1e-13)
    np.testing.assert_allclose(out["Dk"][j], D12, rtol=1e-13)


def test_transport_physical_values(h2mech, h2oracle):
    """Published values pin units: mu(N2, 300 K) = 1.79e-5 Pa s,
    mu(O2, 300 K) = 2.07e-5 Pa s, D(H2-N2, 298 K, 1 atm) = 7.8e-5 m^2/s."""
    sp = h2mech.species
    T, p = np.array([300.0]), np.array([PATM])
    for name, ref in (("N2", 1.79e-5), ("O2", 2.07e-5)):
        Y = np.zeros((9, 1))
        Y[sp.index(name)] = 1.0
        assert h2oracle.mix_props(T, p, Y)["mu"][0] == pytest.approx(ref, rel=0.03), name
    Y = np.zeros((9, 1))
    Y[sp.index("H2")], Y[sp.index("N2")] = 0.5, 0.5
    D = h2oracle.mix_props(np.array([298.0]), p, Y)["Dk"][sp.index("H2"), 0]
    assert D == pytest.approx(7.8e-5, rel=0.05)


def test_wilke_mixture_rules(h2mech, h2oracle):
    """phi_ii = 1, so a mixture of two copies of one gas reduces to the pure
    gas; the mixture viscosity lies between the pure values for a binary."""
    sp = h2mech.species
    T, p = np.array([1200.0]), np.array([PATM])
    k = sp.index("O2")
    pure = {}
    for name in ("O2", "N2"):
        Y = np.zeros((9, 1))
        Y[sp.index(name)] = 1.0
        pure[name] = h2oracle.mix_props(T, p, Y)["mu"][0]
    Y = np.zeros((9, 1))
    Y[k], Y[sp.index("N2")] = 0.4, 0.6
    mu = h2oracle.mix_props(T, p, Y)["mu"][0]
    assert min(pure.values()) < mu < max(pure.values())
    # Wilke's formula evaluated here from the pure-species viscosities
    W = h2mech.tables()["W"]
    idx = [k, sp.index("N2")]
    x = Y[idx, 0] / W[idx]
    x /= x.sum()
    m = np.array([pure["O2"], pure["N2"]])
    w = W[idx]
    phi = np.array([[(1 + np.sqrt(m[a] / m[b] * np.sqrt(w[b] / w[a]))) ** 2 /
                     np.sqrt(8 * (1 + w[a] / w[b])) for b in range(2)] for a in range(2)])
    mix = sum(x[a] * m[a] / (phi[a] @ x) for a in range(2))
    assert mu == pytest.approx(mix, rel=1e-13)


def test_explicit_euler_restatement_converges(h2mech, h2oracle):
    """The reference GPU integrator's restatement (explicit Euler, step
    doubling) approaches a tight implicit solution as rtol tightens."""
    from openhurricane_amd import _lib
    from openhurricane_amd.sampler import h2air_states

    n = 8
    st = h2air_states(h2mech.species, h2mech.tables()["W"], n, seed=11)
    dt = np.full(n, 1e-6)
    ref = _lib.default_opts(atol=1e-16, rtol=1e-9, atol_inc=1e-16, rtol_inc=1e-9)
    Tr, Yr, _, _ = h2oracle.solve_update(ref, st["T"], st["p"], st["rho"], st["Y"], dt, 1.0,
                                         subDt=np.full(n, 1e-13))
    errs = []
    for rtol in (1e-3, 1e-5, 1e-7):
        o = _lib.default_opts(atol=1e-14, rtol=rtol, integrator=1)
        T, Y, _, s = h2oracle.solve_update(o, st["T"], st["p"], st["rho"], st["Y"], dt, 1.0)
        assert s["newton_iters"] == 0 and s["rhs_evals"] == s["attempts"] + s["steps"]
        np.testing.assert_allclose(Y.sum(0), 1.0, rtol=1e-13)
        errs.append(np.max(np.abs(Y - Yr) / np.maximum(Yr, 1e-3)))
    assert errs[0] > errs[1] > errs[2], errs


def test_trapezoidal_restatement_converges_at_second_order(h2mech, h2oracle):
    """ImplicitTrapezoidal (integrator = 2) approaches the tight Radau
    solution as the step tolerance tightens, and ahead of ImplicitEuler at
    the same tolerance (second against first order)."""
    from scipy.integrate import solve_ivp

    from openhurricane_amd import _lib
    from openhurricane_amd.sampler import h2air_states

    W = h2mech.tables()["W"]
    n, dt = 6, 1e-6
    st = h2air_states(h2mech.species, W, n, seed=11, T_range=(1600, 2500))
    refT, refY = np.zeros(n), np.zeros((9, n))
    for i in range(n):
        y0 = np.concatenate([st["rho"][i] * st["Y"][:, i] / W, [st["T"][i]]])
        p = st["p"][i]
        sol = solve_ivp(lambda t, y: h2oracle.dydt(y, p), (0, dt), y0, method="Radau",
                        rtol=1e-12, atol=1e-22,
                        jac=lambda t, y: h2oracle.jacobian(y, p, exact_dkp=True)[1])
        y = sol.y[:, -1]
        refT[i] = y[-1]
        refY[:, i] = y[:-1] * W / (y[:-1] * W).sum()
    errs, steps = {}, {}
    for integ in (0, 2):
        errs[integ] = []
        for rtol in (1e-4, 1e-6, 1e-8):
            o = _lib.default_opts(atol=1e-16, rtol=rtol, atol_inc=1e-16, rtol_inc=rtol,
                                  integrator=integ)
            T, Y, _, s = h2oracle.solve_update(o, st["T"], st["p"], st["rho"], st["Y"],
                                               np.full(n, dt), 1.0, subDt=np.full(n, 1e-13))
            errs[integ].append(max(np.max(np.abs(T - refT) / refT),
                                   np.max(np.abs(Y - refY) / np.maximum(refY, 1e-3))))
            steps[(integ, rtol)] = s["steps"]
    e = errs[2]
    print("trapezoidal", e, "implicit Euler", errs[0], steps)
    assert e[0] > e[1] > e[2], e
    assert e[2] < 3e-4, e
    # fewer steps than backward Euler for the same step tolerance
    assert steps[(2, 1e-8)] < steps[(0, 1e-8)]


def test_bdf23_restatement(h2mech, h2oracle):
    """BDF23 (integrator = 3, order 2): approaches the tight Radau solution as
    the step tolerance tightens within the range where the reference's
    controller works, is at least as accurate as ImplicitEuler at the same
    tolerance, and -- as the reference does -- gives up with a step-size
    underflow at tolerances too tight for its error estimate."""
    from scipy.integrate import solve_ivp

    from openhurricane_amd import _lib
    from openhurricane_amd.sampler import h2air_states

    W = h2mech.tables()["W"]
    n, dt = 6, 1e-6
    st = h2air_states(h2mech.species, W, n, seed=11, T_range=(1600, 2500))
    refT, refY = np.zeros(n), np.zeros((9, n))
    for i in range(n):
        y0 = np.concatenate([st["rho"][i] * st["Y"][:, i] / W, [st["T"][i]]])
        p = st["p"][i]
        sol = solve_ivp(lambda t, y: h2oracle.dydt(y, p), (0, dt), y0, method="Radau",
                        rtol=1e-12, atol=1e-22,
                        jac=lambda t, y: h2oracle.jacobian(y, p, exact_dkp=True)[1])
        y = sol.y[:, -1]
        refT[i] = y[-1]
        refY[:, i] = y[:-1] * W / (y[:-1] * W).sum()

    def err(integ, rtol):
        o = _lib.default_opts(atol=1e-16, rtol=rtol, atol_inc=1e-16, rtol_inc=rtol,
                              integrator=integ)
        T, Y, _, _ = h2oracle.solve_update(o, st["T"], st["p"], st["rho"], st["Y"],
                                           np.full(n, dt), 1.0, subDt=np.full(n, 1e-13))
        return max(np.max(np.abs(T - refT) / refT), np.max(np.abs(Y - refY) / np.maximum(refY, 1e-3)))

    e3, e4 = err(3, 1e-3), err(3, 1e-4)
    print("BDF23", e3, e4, "ImplicitEuler", err(0, 1e-4))
    assert e3 > 2 * e4 and e4 < 0.02
    assert e4 <= 1.1 * err(0, 1e-4)
    o = _lib.default_opts(atol=1e-16, rtol=1e-8, atol_inc=1e-16, rtol_inc=1e-8, integrator=3)
    with pytest.raises(RuntimeError):
        h2oracle.solve_update(o, st["T"], st["p"], st["rho"], st["Y"], np.full(n, dt), 1.0,
                              subDt=np.full(n, 1e-13))
