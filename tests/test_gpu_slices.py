"""Host staging: memory-driven slicing and asynchronous device-pointer calls.

* Host-pointer calls larger than ohc_opts.max_slice_cells run slice by slice
  (the reference's updateChemSourceBySlice / updatePropertiesBySlice,
  turbulentSpeciesSolverCUDATran.cpp:196-222, :525-533): every output equals
  the unsliced call bit for bit, counters sum to the same totals, and the
  padding of species blocks with ld > n is left alone.
* A device-pointer call returns before its kernels finish, so host work
  overlaps the GPU advance.
"""
import time

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]

TOL = dict(atol=1e-10, rtol=1e-6, atol_inc=1e-10, rtol_inc=1e-6)


@pytest.mark.parametrize("kind", ["h2air", "syn32"])
def test_sliced_host_calls_equal_unsliced(kind):
    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource, Mechanism
    from openhurricane_amd.sampler import h2air_states, synthetic_workload

    if kind == "h2air":
        n, m = 5000, 1200
        mech = Mechanism.h2air()
        st = h2air_states(mech.species, mech.tables()["W"], n, seed=61)
    else:
        n, m = 700, 160
        mech = Mechanism.synthetic(32, 150, seed=1)
        st = synthetic_workload(mech.species, mech.tables()["W"], n, seed=62, front=0.2)
    print(f"slices {kind}: {n} cells, {m} per slice", flush=True)
    whole = ChemistrySource(mech, opts=_lib.default_opts(**TOL))
    sliced = ChemistrySource(mech, opts=_lib.default_opts(**TOL, max_slice_cells=m))
    dt = np.full(n, 1e-6)
    outs = []
    for cs in (whole, sliced):
        T, Y = st["T"].copy(), st["Y"].copy()
        sub, stats = cs.solve_update(T, st["p"], st["rho"], Y, dt, 1.0)
        R, sub2, _ = cs.solve(st["T"], st["p"], st["rho"], st["Y"], dt, 1.0)
        Rs, D = cs.source_terms(st["T"], st["p"], st["rho"], st["Y"], diag=True)
        Rk, Dk = cs.source_terms_kim(st["T"], st["p"], st["rho"], st["Y"], fp32=True)
        props = cs.mix_props(T, st["p"], Y, hk=True)
        Th = cs.T_from_h(props["ha"], st["p"], Y, np.full(n, 1500.0))
        outs.append([T, Y, sub, R, sub2, Rs, D, Rk, Dk, *props.values(), Th])
        outs[-1].append(np.array([stats[k] for k in ("cells", "steps", "attempts",
                                                       "newton_iters", "rhs_evals",
                                                       "max_cell_steps", "dt_min")]))
    for i, (a, b) in enumerate(zip(*outs)):
        assert np.array_equal(a, b), i


def test_sliced_padded_leading_dimension():
    import ctypes as C

    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource, Mechanism
    from openhurricane_amd.sampler import h2air_states

    n, ld, m = 1000, 1031, 256
    mech = Mechanism.h2air()
    st = h2air_states(mech.species, mech.tables()["W"], n, seed=63)
    ns = mech.ns
    ref = ChemistrySource(mech, opts=_lib.default_opts(**TOL))
    T0, Y0 = st["T"].copy(), st["Y"].copy()
    ref.solve_update(T0, st["p"], st["rho"], Y0, np.full(n, 1e-6), 1.0)
    cs = ChemistrySource(mech, opts=_lib.default_opts(**TOL, max_slice_cells=m))
    Y = np.full((ns, ld), 9.0)
    Y[:, :n] = st["Y"]
    T = st["T"].copy()
    sub = np.zeros(n)
    dt = np.full(n, 1e-6)  # held by a name: the call reads it after the arguments are built
    rc = _lib.lib().ohc_chem_solve_update(cs._h, n, cs._p(T), cs._p(st["p"]), cs._p(st["rho"]),
                                          C.c_void_p(Y.ctypes.data), ld, cs._p(dt), 1.0,
                                          cs._p(sub), None)
    _lib.check(rc, cs._h)
    assert np.array_equal(T, T0) and np.array_equal(Y[:, :n], Y0)
    assert np.all(Y[:, n:] == 9.0)


def test_device_call_overlaps_host_work():
    """A device-pointer solveUpdate is enqueued and returns; the host then
    does its own work while the GPU integrates."""
    import torch

    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource, Mechanism
    from openhurricane_amd.sampler import synthetic_workload

    n = 1 << 16
    mech = Mechanism.synthetic(32, 150, seed=1)
    st = synthetic_workload(mech.species, mech.tables()["W"], n, seed=64, front=0.2)
    cs = ChemistrySource(mech, opts=_lib.default_opts(**TOL, device_pointers=1))
    g = {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda:0") for k, v in st.items()
         if k != "kind"}
    dt = torch.full((n,), 1e-6, dtype=torch.float64, device="cuda:0")
    sub = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    T, Y = g["T"].clone(), g["Y"].clone()
    cs.solve_update(T, g["p"], g["rho"], Y, dt, 1.0, subDt=sub, stats=False)  # warm-up
    torch.cuda.synchronize()
    T, Y = g["T"].clone(), g["Y"].clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cs.solve_update(T, g["p"], g["rho"], Y, dt, 1.0, subDt=sub, stats=False)
    t_call = time.perf_counter() - t0
    host = float(np.sum(np.sqrt(np.arange(2_000_000, dtype=np.float64))))  # host work
    t_host = time.perf_counter() - t0
    torch.cuda.synchronize()
    t_all = time.perf_counter() - t0
    print(f"call returned after {t_call * 1e3:.2f} ms, host work done at {t_host * 1e3:.2f} ms, "
          f"GPU done at {t_all * 1e3:.2f} ms", flush=True)
    assert host > 0 and t_call < 0.25 * t_all
