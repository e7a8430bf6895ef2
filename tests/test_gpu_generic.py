"""GPU parity of the generic wavefront-per-cell kernels (run-time mechanism
tables) against the FP64 oracle, through the C-ABI.

Cases: the H2-air mechanism forced onto the generic path (so the two device
paths and the oracle agree on the same cells), and the synthetic 32-species /
150-reaction mechanism (third-body, Lindemann / Troe / SRI fall-off), both
reactor types and both Jacobian conventions.  Tolerances as for the lane
kernels: RHS / Jacobian 1e-9 of the row scale; integrated states
|dT|/T <= 1e-6, |dY| <= 1e-6 max(Y, 1e-3)."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]


def _opts(**kw):
    from openhurricane_amd import _lib

    base = dict(atol=1e-10, rtol=1e-6, atol_inc=1e-10, rtol_inc=1e-6, kernel=2)
    base.update(kw)
    return _lib.default_opts(**base)


@pytest.fixture(scope="module")
def syn32():
    from openhurricane_amd.chemistry import Mechanism

    return Mechanism.synthetic(32, 150, seed=1)


@pytest.fixture(scope="module")
def syn32_oracle(syn32):
    from oracle import Oracle

    return Oracle(syn32)


def _ode_states(st, W):
    c = st["rho"] * st["Y"] / W[:, None]
    return np.ascontiguousarray(np.vstack([c, st["T"][None, :]]))


def _check_eval(cs, orc, y, p, flow, exact):
    f, J = cs.eval_jacobian(y, p)
    N = y.shape[0]
    for i in range(y.shape[1]):
        fo, Jo = orc.jacobian(y[:, i], p[i], flow=flow, exact_dkp=exact)
        scale = np.max(np.abs(fo[:-1])) + 1e-300
        assert np.all(np.abs(f[:-1, i] - fo[:-1]) <= 1e-9 * scale), i
        assert abs(f[-1, i] - fo[-1]) <= 1e-9 * abs(fo[-1]) + 1e-9 * scale
        row = np.max(np.abs(Jo), axis=1, keepdims=True) + 1e-300
        err = np.abs(J[:, i].reshape(N, N) - Jo) / row
        assert err.max() <= 1e-9, (i, err.max(), np.unravel_index(err.argmax(), err.shape))


def _check_states(T, Y, To, Yo):
    dT = np.abs(T - To) / To
    dY = np.abs(Y - Yo) / np.maximum(np.abs(Yo), 1e-3)
    print("max |dT|/T", dT.max(), "max dY", dY.max())
    assert dT.max() <= 1e-6
    assert dY.max() <= 1e-6


def test_generic_h2air_matches_oracle(h2mech, h2oracle):
    from openhurricane_amd.chemistry import ChemistrySource
    from openhurricane_amd.sampler import h2air_states

    o = _opts()
    cs = ChemistrySource(h2mech, opts=o)
    assert cs.kernel_name.startswith("wave:"), cs.kernel_name
    t = h2mech.tables()
    st = h2air_states(h2mech.species, t["W"], 512, seed=7)
    y = _ode_states(st, t["W"])
    _check_eval(cs, h2oracle, y[:, :128].copy(), np.ascontiguousarray(st["p"][:128]), 0, False)
    dt = np.full(512, 1e-6)
    To, Yo, sub_o, s_o = h2oracle.solve_update(o, st["T"], st["p"], st["rho"], st["Y"], dt, 1.0,
                                               nthreads=16)
    T, Y = st["T"].copy(), st["Y"].copy()
    sub, s = cs.solve_update(T, st["p"], st["rho"], Y, dt, 1.0)
    _check_states(T, Y, To, Yo)
    assert s["cells"] == 512
    assert abs(s["steps"] - s_o["steps"]) <= 0.01 * s_o["steps"]
    np.testing.assert_allclose(sub, sub_o, rtol=1e-3)


@pytest.mark.parametrize("flow,exact", [(0, 0), (0, 1), (1, 1)])
def test_generic_synthetic_eval_matches_oracle(syn32, syn32_oracle, flow, exact):
    from openhurricane_amd.chemistry import ChemistrySource
    from openhurricane_amd.sampler import synthetic_workload

    cs = ChemistrySource(syn32, opts=_opts(flow_type=flow, jacobian=exact))
    t = syn32.tables()
    st = synthetic_workload(syn32.species, t["W"], 96, seed=5)
    y = _ode_states(st, t["W"])
    # hot states too, so that the fall-off reactions run in their fall-off range
    y[-1, ::3] = np.linspace(1700.0, 2600.0, y[:, ::3].shape[1])
    pr = st["p"] if flow == 0 else st["rho"]
    _check_eval(cs, syn32_oracle, y, np.ascontiguousarray(pr), flow, bool(exact))


@pytest.mark.parametrize("flow,exact,n", [(0, 1, 96), (0, 0, 24), (1, 1, 48)])
def test_generic_synthetic_solve_matches_oracle(syn32, syn32_oracle, flow, exact, n):
    from openhurricane_amd.chemistry import ChemistrySource
    from openhurricane_amd.sampler import synthetic_workload

    o = _opts(flow_type=flow, jacobian=exact, atol_inc=1e-7, rtol_inc=1e-3)
    cs = ChemistrySource(syn32, opts=o)
    t = syn32.tables()
    st = synthetic_workload(syn32.species, t["W"], n, seed=11 + flow)
    dt = np.full(n, 1e-6)
    To, Yo, sub_o, s_o = syn32_oracle.solve_update(o, st["T"], st["p"], st["rho"], st["Y"], dt,
                                                   1.0, nthreads=16)
    T, Y = st["T"].copy(), st["Y"].copy()
    sub, s = cs.solve_update(T, st["p"], st["rho"], Y, dt, 1.0)
    print("gpu", s, "oracle", s_o)
    _check_states(T, Y, To, Yo)
    assert abs(s["newton_iters"] - s_o["newton_iters"]) <= 0.02 * s_o["newton_iters"]
    # element moles per unit mass drift only through the clipping of negative
    # concentrations in ciToYi (loose step tolerances): the same drift as the oracle's
    E, W = t["elem"], t["W"]
    el0 = (st["Y"] / W[:, None]).T @ E
    drift = ((Y / W[:, None]).T @ E - el0) / np.maximum(np.abs(el0), 1e-12)
    drift_o = ((Yo / W[:, None]).T @ E - el0) / np.maximum(np.abs(el0), 1e-12)
    np.testing.assert_allclose(drift, drift_o, rtol=0, atol=1e-9)
    # source-term mode on the same cells
    R, sub2, s2 = cs.solve(st["T"], st["p"], st["rho"], st["Y"], dt, 1.0)
    Ro, _, _, _ = syn32_oracle.solve(o, st["T"], st["p"], st["rho"], st["Y"], dt, 1.0, nthreads=16)
    scale = np.max(np.abs(Ro), axis=0) + 1e-300
    assert np.max(np.abs(R - Ro) / scale) <= 1e-5


def test_generic_burnt_states_match_oracle(syn32, syn32_oracle):
    """The benchmark's burnt cells: fresh hot mixtures (1800-2600 K)
    integrated over 1e-2 s (large sub-steps, where I - h J needs pivoting
    away from the diagonal), then the burnt states over one 1e-6 s flow
    step, with the reference Jacobian: states and Newton counts of the
    oracle."""
    from openhurricane_amd.chemistry import ChemistrySource
    from openhurricane_amd.sampler import fresh_states, mixture_density

    t = syn32.tables()
    n = 64
    b = fresh_states(syn32.species, t["W"], n, np.random.Generator(np.random.PCG64(81)),
                     (1800.0, 2600.0))
    for exact, span in ((1, 1e-2), (0, 1e-6)):
        o = _opts(jacobian=exact, atol_inc=1e-7, rtol_inc=1e-3)
        cs = ChemistrySource(syn32, opts=o)
        dt = np.full(n, span)
        print(f"burnt states: {n} cells over {span} s, jacobian {exact}", flush=True)
        To, Yo, _, s_o = syn32_oracle.solve_update(o, b["T"], b["p"], b["rho"], b["Y"], dt, 1.0,
                                                   nthreads=16)
        T, Y = b["T"].copy(), b["Y"].copy()
        _, s = cs.solve_update(T, b["p"], b["rho"], Y, dt, 1.0)
        print("gpu", s, "oracle", s_o, flush=True)
        _check_states(T, Y, To, Yo)
        assert abs(s["newton_iters"] - s_o["newton_iters"]) <= 0.02 * s_o["newton_iters"]
        b = {"T": To, "p": b["p"], "rho": mixture_density(b["p"], To, Yo, t["W"]), "Y": Yo}


def test_generic_schedule_is_bitwise_neutral(syn32):
    from openhurricane_amd.chemistry import ChemistrySource
    from openhurricane_amd.sampler import synthetic_workload

    t = syn32.tables()
    n = 256
    st = synthetic_workload(syn32.species, t["W"], n, seed=4)
    dt = np.full(n, 1e-6)
    outs = []
    cs1 = ChemistrySource(syn32, opts=_opts(jacobian=1, schedule=1, atol_inc=1e-7, rtol_inc=1e-3))
    cs0 = ChemistrySource(syn32, opts=_opts(jacobian=1, schedule=0, atol_inc=1e-7, rtol_inc=1e-3))
    for c in (cs0, cs1, cs1):
        T, Y = st["T"].copy(), st["Y"].copy()
        sub, s = c.solve_update(T, st["p"], st["rho"], Y, dt, 1.0)
        outs.append((T, Y, sub))
    for T, Y, sub in outs[1:]:
        np.testing.assert_array_equal(T, outs[0][0])
        np.testing.assert_array_equal(Y, outs[0][1])
        np.testing.assert_array_equal(sub, outs[0][2])


@pytest.mark.parametrize("without_last", [0, 1])
def test_generic_source_terms_match_oracle(syn32, syn32_oracle, without_last):
    from openhurricane_amd.chemistry import ChemistrySource
    from openhurricane_amd.sampler import synthetic_workload

    o = _opts()
    cs = ChemistrySource(syn32, opts=o)
    t = syn32.tables()
    st = synthetic_workload(syn32.species, t["W"], 128, seed=13)
    st["T"][::2] = np.linspace(1700.0, 2600.0, 64)
    R, D = cs.source_terms(st["T"], st["p"], st["rho"], st["Y"], diag=True,
                           without_last=bool(without_last))
    Ro, Do = syn32_oracle.chem_source(o, st["T"], st["p"], st["rho"], st["Y"], diag=True,
                                      without_last=bool(without_last))
    assert np.all(np.abs(R - Ro) <= 1e-9 * np.abs(Ro).max(axis=0) + 1e-300)
    nd = 32 - without_last
    sc = np.abs(Do[:nd]).max(axis=0) + 1e-300
    assert np.all(np.abs(D[:nd] - Do[:nd]) <= 1e-9 * sc)


def test_generic_54_species_matches_oracle():
    """The 54-species / 270-reaction mechanism (wave size 55, two wavefronts
    per SIMD): RHS / Jacobian and an integrated step against the oracle."""
    from oracle import Oracle

    from openhurricane_amd.chemistry import ChemistrySource, Mechanism
    from openhurricane_amd.sampler import synthetic_workload

    mech = Mechanism.synthetic(54, 270, seed=2)
    orc = Oracle(mech)
    t = mech.tables()
    o = _opts(jacobian=1, atol_inc=1e-7, rtol_inc=1e-3)
    cs = ChemistrySource(mech, opts=o)
    assert cs.kernel_name == "wave:55", cs.kernel_name
    st = synthetic_workload(mech.species, t["W"], 48, seed=13, front=0.3)
    y = _ode_states(st, t["W"])
    y[-1, ::4] = np.linspace(1700.0, 2600.0, y[:, ::4].shape[1])
    _check_eval(cs, orc, y, np.ascontiguousarray(st["p"]), 0, True)
    dt = np.full(48, 1e-6)
    To, Yo, _, s_o = orc.solve_update(o, st["T"], st["p"], st["rho"], st["Y"], dt, 1.0, nthreads=16)
    T, Y = st["T"].copy(), st["Y"].copy()
    _, s = cs.solve_update(T, st["p"], st["rho"], Y, dt, 1.0)
    print("gpu", s, "oracle", s_o)
    _check_states(T, Y, To, Yo)
    assert abs(s["newton_iters"] - s_o["newton_iters"]) <= 0.02 * s_o["newton_iters"]


@pytest.mark.parametrize("which", ["h2air", "syn32", "syn54"])
def test_generic_props_and_T_from_h_match_oracle(which, h2mech, h2oracle, syn32, syn32_oracle):
    """Mixture thermo / transport (wavefront reductions, pair sums over LDS)
    and T from h (table-driven Newton) on the run-time tables, against the
    oracle with the tolerances of the compiled-mechanism test."""
    from oracle import Oracle

    from openhurricane_amd.chemistry import ChemistrySource, Mechanism
    from openhurricane_amd.sampler import h2air_states, synthetic_workload

    if which == "h2air":
        mech, orc = h2mech, h2oracle
    elif which == "syn32":
        mech, orc = syn32, syn32_oracle
    else:
        mech = Mechanism.synthetic(54, 270, seed=2)
        orc = Oracle(mech)
    t = mech.tables()
    o = _opts()
    cs = ChemistrySource(mech, opts=o)
    assert cs.kernel_name.startswith("wave:"), cs.kernel_name
    n = 512
    if which == "h2air":
        st = h2air_states(mech.species, t["W"], n, seed=19)
    else:
        st = synthetic_workload(mech.species, t["W"], n, seed=19)
    out = cs.mix_props(st["T"], st["p"], st["Y"], hk=True)
    ref = orc.mix_props(st["T"], st["p"], st["Y"], hk=True)
    for k in ["mu", "kappa", "cp", "ha"]:
        np.testing.assert_allclose(out[k], ref[k], rtol=1e-12, atol=0, err_msg=k)
    np.testing.assert_allclose(out["Dk"], ref["Dk"], rtol=1e-11, atol=1e-300)
    np.testing.assert_allclose(out["hk"], ref["hk"], rtol=1e-12, atol=1e-6)
    T0 = st["T"] + 300.0
    Tg, fl = cs.T_from_h(ref["ha"], st["p"], st["Y"], T0, want_flag=True)
    To, flo = orc.T_from_h(o, ref["ha"], st["p"], st["Y"], T0)
    np.testing.assert_allclose(Tg, To, rtol=1e-12)
    np.testing.assert_array_equal(fl, flo)
    np.testing.assert_allclose(Tg, st["T"], rtol=1e-4)


@pytest.mark.parametrize("which,flow", [("h2air", 0), ("syn32", 0), ("syn32", 1)])
def test_generic_trapezoidal_matches_oracle(which, flow, h2mech, h2oracle, syn32, syn32_oracle):
    """ImplicitTrapezoidal (integrator = 2) on the generic kernels against the
    oracle's restatement: states, step and Newton counts."""
    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource
    from openhurricane_amd.sampler import h2air_states, synthetic_workload

    if which == "h2air":
        mech, orc = h2mech, h2oracle
        o = _lib.default_opts(atol=1e-10, rtol=1e-6, atol_inc=1e-10, rtol_inc=1e-6, integrator=2,
                              flow_type=flow, kernel=2)
        n = 512
        st = h2air_states(mech.species, mech.tables()["W"], n, seed=23)
    else:
        mech, orc = syn32, syn32_oracle
        o = _opts(integrator=2, jacobian=1, atol_inc=1e-7, rtol_inc=1e-3, flow_type=flow)
        n = 96
        st = synthetic_workload(mech.species, mech.tables()["W"], n, seed=29)
    cs = ChemistrySource(mech, opts=o)
    assert cs.kernel_name.startswith("wave:"), cs.kernel_name
    dt = np.full(n, 1e-6)
    To, Yo, sub_o, s_o = orc.solve_update(o, st["T"], st["p"], st["rho"], st["Y"], dt, 1.0,
                                          nthreads=16)
    T, Y = st["T"].copy(), st["Y"].copy()
    sub, s = cs.solve_update(T, st["p"], st["rho"], Y, dt, 1.0)
    print("gpu", s, "oracle", s_o)
    _check_states(T, Y, To, Yo)
    assert abs(s["steps"] - s_o["steps"]) <= 0.01 * s_o["steps"]
    assert abs(s["newton_iters"] - s_o["newton_iters"]) <= 0.02 * s_o["newton_iters"]


def _bdf23_case(which, flow, dt, h2mech, h2oracle, syn32, syn32_oracle, jac=1):
    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource
    from openhurricane_amd.sampler import h2air_states, synthetic_workload

    if which == "h2air":
        mech, orc = h2mech, h2oracle
        o = _lib.default_opts(atol=1e-10, rtol=1e-6, atol_inc=1e-8, rtol_inc=1e-3, integrator=3,
                              flow_type=flow, kernel=2)
        n = 512
        st = h2air_states(mech.species, mech.tables()["W"], n, seed=23)
    else:
        mech, orc = syn32, syn32_oracle
        o = _opts(integrator=3, jacobian=jac, atol_inc=1e-8, rtol_inc=1e-3, flow_type=flow)
        n = 96
        st = synthetic_workload(mech.species, mech.tables()["W"], n, seed=29)
    cs = ChemistrySource(mech, opts=o)
    assert cs.kernel_name.startswith("wave:"), cs.kernel_name
    dtv = np.full(n, dt)
    To, Yo, sub_o, s_o = orc.solve_update(o, st["T"], st["p"], st["rho"], st["Y"], dtv, 1.0,
                                          nthreads=16)
    T, Y = st["T"].copy(), st["Y"].copy()
    sub, s = cs.solve_update(T, st["p"], st["rho"], Y, dtv, 1.0)
    print("gpu", s, "oracle", s_o)
    return T, Y, s, To, Yo, s_o


@pytest.mark.parametrize("flow,jac", [(0, 1), (1, 1), (0, 0)])
def test_generic_bdf23_matches_oracle(flow, jac, h2mech, h2oracle, syn32, syn32_oracle):
    """BDF23 (integrator = 3, order 2) on the generic kernels against the
    oracle's restatement, at the reference's BDF23 step tolerances
    (atol 1e-8, rtol 1e-3), on the synthetic 32-species mechanism over 1 us:
    states, step and Newton counts.  Here the algorithm is well conditioned:
    perturbing the oracle's input temperatures by 1e-14 moves no final
    temperature by more than 3.3e-13 with the reference's Jacobian
    (jacobian = 1) and 3.4e-8 with the exact one, and the step counts by at
    most 5 of 8506."""
    T, Y, s, To, Yo, s_o = _bdf23_case("syn32", flow, 1e-6, h2mech, h2oracle, syn32,
                                       syn32_oracle, jac)
    _check_states(T, Y, To, Yo)
    assert abs(s["steps"] - s_o["steps"]) <= 0.01 * s_o["steps"]
    assert abs(s["newton_iters"] - s_o["newton_iters"]) <= 0.02 * s_o["newton_iters"]


@pytest.mark.parametrize("dt", [1e-7, 1e-6])
def test_generic_bdf23_sensitive_regime(dt, h2mech, h2oracle, syn32, syn32_oracle):
    """H2-air at the BDF23 tolerances is a regime where the reference's step
    control is itself chaotic: its error estimate stops falling with the step
    size, and whether an attempt is accepted ("error grew from the last
    attempt's and is below 5") turns on rounding.  On the oracle, perturbing
    the input temperatures of these 512 cells by 1e-15 .. 1e-13 changes up
    to 7 cells by up to 8.6e-5 in Y over 0.1 us, and over 1 us changes 76
    cells by up to 1.1e-4 in T and 5e-3 in Y and makes one fail with a
    step-size underflow.  The GPU run must stay within that envelope: most
    cells identical, the others within the step tolerance, step counts
    within 2 %."""
    T, Y, s, To, Yo, s_o = _bdf23_case("h2air", 0, dt, h2mech, h2oracle, syn32, syn32_oracle)
    dT = np.abs(T - To) / To
    print("cells off by > 1e-9:", int((dT > 1e-9).sum()), "max", dT.max())
    assert (dT > 1e-9).sum() <= 128
    assert dT.max() <= 1e-3
    assert np.all(np.abs(Y - Yo) <= 1e-2 * np.maximum(np.abs(Yo), 1e-3))
    assert abs(s["steps"] - s_o["steps"]) <= 0.02 * s_o["steps"]


def test_bdf23_options_are_checked(h2mech):
    """BDF23 runs on the generic kernels in FP64 only: the lane kernel and the
    FP32 split are rejected at context creation."""
    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource

    with pytest.raises(_lib.OhcError):
        ChemistrySource(h2mech, opts=_lib.default_opts(integrator=3, kernel=1))
    with pytest.raises(_lib.OhcError):
        ChemistrySource(h2mech, opts=_lib.default_opts(integrator=3, precision=1))
    cs = ChemistrySource(h2mech, opts=_lib.default_opts(integrator=3))
    assert cs.kernel_name.startswith("wave:"), cs.kernel_name


@pytest.mark.parametrize("which", ["h2air_lane", "h2air", "syn32", "syn54"])
def test_energy_forms_match_oracle(which, h2mech, h2oracle, syn32, syn32_oracle):
    """thermoList::cv_p / ea_p and TEa (T from internal energy) on the
    compiled H2-air kernels and on the run-time-table kernels, against the
    oracle: cv, ea to 1e-12; T(e) to 1e-12 with identical limiter flags."""
    from oracle import Oracle

    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource, Mechanism
    from openhurricane_amd.sampler import h2air_states, synthetic_workload

    if which.startswith("h2air"):
        mech, orc = h2mech, h2oracle
    elif which == "syn32":
        mech, orc = syn32, syn32_oracle
    else:
        mech = Mechanism.synthetic(54, 270, seed=2)
        orc = Oracle(mech)
    t = mech.tables()
    o = _opts() if which != "h2air_lane" else _lib.default_opts()
    cs = ChemistrySource(mech, opts=o)
    assert cs.kernel_name.startswith("lane:" if which == "h2air_lane" else "wave:"), cs.kernel_name
    n = 512
    if which.startswith("h2air"):
        st = h2air_states(mech.species, t["W"], n, seed=29, T_range=(300.0, 3500.0))
    else:
        st = synthetic_workload(mech.species, t["W"], n, seed=29)
    Y = np.ascontiguousarray(st["Y"])
    out = cs.mix_thermo(st["T"], st["p"], Y)
    ref = orc.mix_thermo(st["T"], st["p"], Y)
    for k in ("cp", "cv", "ha", "ea"):
        np.testing.assert_allclose(out[k], ref[k], rtol=1e-12, atol=1e-9 * np.abs(ref[k]).max(),
                                   err_msg=k)
    T0 = st["T"] + 300.0
    Tg, fl = cs.T_from_e(ref["ea"], st["p"], Y, T0, want_flag=True)
    To, flo = orc.T_from_e(o, ref["ea"], st["p"], Y, T0)
    np.testing.assert_allclose(Tg, To, rtol=1e-12)
    np.testing.assert_array_equal(fl, flo)
    np.testing.assert_allclose(Tg, st["T"], rtol=1e-4)


@pytest.mark.parametrize("kind", ["h2air", "syn32"])
def test_generic_split_precision_against_oracle(kind):
    """The FP32 split on the generic (run-time table) kernels: errors against
    the FP64 oracle per output, bounded on the non-runaway cells."""
    from oracle import Oracle

    from openhurricane_amd.chemistry import Mechanism
    from openhurricane_amd.sampler import h2air_states, synthetic_workload
    from test_gpu_parity import split_precision_check

    if kind == "h2air":
        mech = Mechanism.h2air()
        st = h2air_states(mech.species, mech.tables()["W"], 1024, seed=71)
        tol = dict(atol=1e-10, rtol=1e-6, atol_inc=1e-10, rtol_inc=1e-6, kernel=2)
        n = 1024
    else:
        mech = Mechanism.synthetic(32, 150, seed=1)
        st = synthetic_workload(mech.species, mech.tables()["W"], 256, seed=72, front=0.2)
        tol = dict(atol=1e-10, rtol=1e-6, atol_inc=1e-7, rtol_inc=1e-3, jacobian=1)
        n = 256
    split_precision_check(mech, Oracle(mech), st, n, tol, "wave:")
