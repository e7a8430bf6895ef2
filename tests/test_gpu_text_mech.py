"""A mechanism read from CHEMKIN text with LOW/TROE (3 and 4 parameters),
SRI, HIGH (chemically activated), REV, DUP and FORD lines, run on the
generic GPU kernels against the oracle: RHS and Jacobian to 1e-9 of the row
scale, one ImplicitEuler advance to |dT|/T, |dY| <= 1e-6.  The oracle's rates
on this text are pinned to the textbook forms in test_reader.py."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]


def _states(mech, n, seed):
    """H2 / O2 / N2 (named LOWN2 here) / H2O mixtures carrying radicals at
    1e-6 .. 1e-3 mass fraction, 900 .. 2300 K, 1 and 20 bar."""
    rng = np.random.default_rng(seed)
    W = mech.tables()["W"]
    sp = mech.species
    Y = np.zeros((len(W), n))
    for k in ("O", "H", "OH", "HO2", "H2O2"):
        Y[sp.index(k)] = 10.0 ** rng.uniform(-6, -3, n)
    Y[sp.index("H2")] = rng.uniform(0.005, 0.05, n)
    Y[sp.index("O2")] = rng.uniform(0.1, 0.23, n)
    Y[sp.index("H2O")] = rng.uniform(0.0, 0.15, n)
    Y[sp.index("LOWN2")] = 1.0 - Y.sum(axis=0)
    T = rng.uniform(900.0, 2300.0, n)
    p = rng.choice([1e5, 2e6], n)
    Ru = 8314.462618
    rho = p / (Ru * T * np.sum(Y / W[:, None], axis=0))
    return {"T": T, "p": p, "rho": rho, "Y": np.ascontiguousarray(Y)}


@pytest.mark.parametrize("exact", [0, 1])
def test_text_mechanism_on_the_generic_kernel(tmp_path, exact):
    from oracle import Oracle
    from test_reader import ONE_LINE, _load

    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource

    mech = _load(tmp_path, ONE_LINE, "gpu")
    orc = Oracle(mech)
    o = _lib.default_opts(atol=1e-10, rtol=1e-6, atol_inc=1e-10, rtol_inc=1e-6, kernel=2,
                          jacobian=exact)
    cs = ChemistrySource(mech, opts=o)
    print(f"text mechanism: {mech.ns} species, {mech.nr} reactions, kernel {cs.kernel_name}, "
          f"jacobian {exact}", flush=True)
    assert cs.kernel_name.startswith("wave:")
    n = 256
    st = _states(mech, n, seed=11 + exact)
    W = mech.tables()["W"]
    y = np.ascontiguousarray(np.vstack([st["rho"] * st["Y"] / W[:, None], st["T"][None, :]]))
    f, J = cs.eval_jacobian(y, st["p"])
    N = y.shape[0]
    for i in range(n):
        fo, Jo = orc.jacobian(y[:, i], st["p"][i], flow=0, exact_dkp=bool(exact))
        scale = np.max(np.abs(fo[:-1])) + 1e-300
        assert np.all(np.abs(f[:-1, i] - fo[:-1]) <= 1e-9 * scale), i
        row = np.max(np.abs(Jo), axis=1, keepdims=True) + 1e-300
        assert np.max(np.abs(J[:, i].reshape(N, N) - Jo) / row) <= 1e-9, i
    dt = np.full(n, 1e-6)
    To, Yo, _, s_o = orc.solve_update(o, st["T"], st["p"], st["rho"], st["Y"], dt, 1.0,
                                      nthreads=16)
    T, Y = st["T"].copy(), st["Y"].copy()
    _, s = cs.solve_update(T, st["p"], st["rho"], Y, dt, 1.0)
    dT = np.max(np.abs(T - To) / To)
    dY = np.max(np.abs(Y - Yo) / np.maximum(np.abs(Yo), 1e-3))
    print(f"max |dT|/T {dT:.2e}, max dY {dY:.2e}, steps {s['steps']} vs {s_o['steps']}")
    assert dT <= 1e-6 and dY <= 1e-6
    assert abs(s["steps"] - s_o["steps"]) <= 0.01 * s_o["steps"] + 2
