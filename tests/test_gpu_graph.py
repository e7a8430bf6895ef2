"""HIP graph capture of the device-pointer chemistry advance.

A caller that keeps its fields on the GPU can capture
ohc_chem_solve_update (no statistics requested, so nothing synchronises) into
a graph and replay it every flow step.  The captured work is the control-word
reset kernel and the integrator kernel -- no memset nodes, no host copies.
Each replay, with the program's own eager work and synchronisations between
replays, must give the eager result bit for bit.  With the longest-first
schedule on (the benchmarked configuration) the graph also holds the
counting-sort kernels that order the next call; the schedule buffers are
allocated by the eager call before the capture (same n).
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]


def _run(mech, states, kernel, n, replays=3, schedule=0):
    import torch

    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource

    dev = torch.device("cuda:0")
    opts = _lib.default_opts(atol=1e-10, rtol=1e-6, atol_inc=1e-7, rtol_inc=1e-3, jacobian=1,
                             device_pointers=1, schedule=schedule, kernel=kernel)
    cs = ChemistrySource(mech, opts=opts)
    g = {k: torch.from_numpy(np.ascontiguousarray(v[..., :n])).to(dev)
         for k, v in states.items() if k != "kind"}
    dt = torch.full((n,), 1e-6, dtype=torch.float64, device=dev)
    sub0 = torch.zeros(n, dtype=torch.float64, device=dev)
    # eager reference
    Te, Ye, se = g["T"].clone(), g["Y"].clone(), sub0.clone()
    cs.solve_update(Te, g["p"], g["rho"], Ye, dt, 1.0, subDt=se, stats=False)
    torch.cuda.synchronize()
    # capture on a side stream the context is bound to before capture starts
    side = torch.cuda.Stream()
    cs.set_stream(side.cuda_stream)
    T, Y, sub = g["T"].clone(), g["Y"].clone(), sub0.clone()
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=side):
        cs.solve_update(T, g["p"], g["rho"], Y, dt, 1.0, subDt=sub, stats=False)
    torch.cuda.synchronize()
    for r in range(replays):
        T.copy_(g["T"])
        Y.copy_(g["Y"])
        sub.copy_(sub0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(T, Te) and torch.equal(Y, Ye) and torch.equal(sub, se), r
        # eager work of the program between replays
        _ = (T * 2.0).sum().item()
    cs.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    return cs.kernel_name


@pytest.mark.parametrize("schedule", [0, 1])
def test_graph_replay_lane_kernel(h2mech, schedule):
    from openhurricane_amd.sampler import h2air_states

    t = h2mech.tables()
    st = h2air_states(h2mech.species, t["W"], 4096, seed=31)
    name = _run(h2mech, st, kernel=1, n=4096, schedule=schedule)
    assert name.startswith("lane:")


@pytest.mark.parametrize("schedule", [0, 1])
def test_graph_replay_generic_kernel(schedule):
    from openhurricane_amd.chemistry import Mechanism
    from openhurricane_amd.sampler import synthetic_workload

    mech = Mechanism.synthetic(32, 150, seed=1)
    t = mech.tables()
    st = synthetic_workload(mech.species, t["W"], 1024, seed=33)
    name = _run(mech, st, kernel=2, n=1024, schedule=schedule)
    assert name.startswith("wave:")


def test_autotune_keeps_results_and_inputs(h2mech):
    """The autotuner times every persistent-grid size on a sample, leaves the
    sample untouched, and the advance with the chosen grid gives bit for bit
    the result of the default grid (each cell is integrated by one lane or
    wavefront whatever the grid)."""
    import torch

    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource
    from openhurricane_amd.sampler import h2air_states

    dev = torch.device("cuda:0")
    t = h2mech.tables()
    n = 16384
    st = h2air_states(h2mech.species, t["W"], n, seed=35)
    g = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in st.items()}
    dt = torch.full((n,), 1e-6, dtype=torch.float64, device=dev)
    opts = _lib.default_opts(atol=1e-10, rtol=1e-6, device_pointers=1)
    cs = ChemistrySource(h2mech, opts=opts)
    T0, Y0 = g["T"].clone(), g["Y"].clone()
    Ta, Ya = g["T"].clone(), g["Y"].clone()
    cs.solve_update(Ta, g["p"], g["rho"], Ya, dt, 1.0)
    w, ms = cs.autotune(g["T"], g["p"], g["rho"], g["Y"], dt, 1.0)
    print("autotuned waves per CU", w, "step", ms, "ms")
    assert w >= 1 and ms > 0
    assert torch.equal(g["T"], T0) and torch.equal(g["Y"], Y0)
    Tb, Yb = g["T"].clone(), g["Y"].clone()
    cs.solve_update(Tb, g["p"], g["rho"], Yb, dt, 1.0)
    assert torch.equal(Ta, Tb) and torch.equal(Ya, Yb)
