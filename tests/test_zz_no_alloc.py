"""ohc_chem_solve_update_noalloc: the solveUpdate entry point that allocates
nothing (queue head, counters and the longest-first schedule in the
library's static device scratch array).

Each call is checked bit for bit against ohc_chem_solve_update on the same
inputs (T, Y, suggested sub-step and counters): two calls in a row, the
second on fewer cells, then a third on the same cells so that the schedule
carried in the scratch array orders it.  Also: a HIP graph capture of the
entry point replays to the eager result, and sizes beyond the scratch limit
and host-pointer contexts are rejected before anything runs.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]

TOL = dict(atol=1e-10, rtol=1e-6, atol_inc=1e-7, rtol_inc=1e-3)


def _states(kind, n):
    from openhurricane_amd.chemistry import Mechanism
    from openhurricane_amd.sampler import h2air_states, synthetic_workload

    if kind == "h2air":
        mech = Mechanism.h2air()
        st = h2air_states(mech.species, mech.tables()["W"], n, seed=11)
    else:
        mech = Mechanism.synthetic(32, 150, seed=1)
        st = synthetic_workload(mech.species, mech.tables()["W"], n, seed=12, front=0.2)
    return mech, st


def _dev(st, n):
    import torch

    return {k: torch.from_numpy(np.ascontiguousarray(v[..., :n])).to("cuda:0")
            for k, v in st.items() if k != "kind"}


def _advance(cs, g, n, noalloc):
    import torch

    T, Y = g["T"][:n].clone(), g["Y"][:, :n].contiguous().clone()
    sub = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    dt = torch.full((n,), 1e-6, dtype=torch.float64, device="cuda:0")
    p, rho = g["p"][:n].contiguous(), g["rho"][:n].contiguous()
    f = cs.solve_update_noalloc if noalloc else cs.solve_update
    _, stats = f(T, p, rho, Y, dt, 1.0, subDt=sub, stats=True)
    torch.cuda.synchronize()
    return T, Y, sub, stats


@pytest.mark.parametrize("kind,n1,n2", [("h2air", 4096, 1500), ("syn32", 384, 200)])
def test_noalloc_matches_solve_update(kind, n1, n2):
    import torch

    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource

    mech, st = _states(kind, n1)
    g = _dev(st, n1)
    opts = _lib.default_opts(**TOL, device_pointers=1, schedule=1)
    ref = ChemistrySource(mech, opts=opts)
    na = ChemistrySource(mech, opts=_lib.default_opts(**TOL, device_pointers=1, schedule=1))
    for step, n in enumerate((n1, n2, n2)):
        print(f"noalloc {kind}: call {step}, {n} cells", flush=True)
        Tr, Yr, sr, str_ = _advance(ref, g, n, noalloc=False)
        Tn, Yn, sn, stn = _advance(na, g, n, noalloc=True)
        assert torch.equal(Tn, Tr), (kind, step)
        assert torch.equal(Yn, Yr), (kind, step)
        assert torch.equal(sn, sr), (kind, step)
        for k in ("cells", "steps", "attempts", "newton_iters", "rhs_evals", "max_cell_steps",
                  "dt_min"):
            assert stn[k] == str_[k], (kind, step, k)
        assert stn["cells"] == n


def test_noalloc_graph_replay():
    import torch

    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource

    n = 2048
    mech, st = _states("h2air", n)
    g = _dev(st, n)
    cs = ChemistrySource(mech, opts=_lib.default_opts(**TOL, device_pointers=1, schedule=1))
    Te, Ye, se, _ = _advance(cs, g, n, noalloc=False)
    dt = torch.full((n,), 1e-6, dtype=torch.float64, device="cuda:0")
    # eager noalloc call first: the capture then holds the scheduled form
    T, Y = g["T"].clone(), g["Y"].clone()
    sub = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    cs.solve_update_noalloc(T, g["p"], g["rho"], Y, dt, 1.0, subDt=sub, stats=False)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    cs.set_stream(side.cuda_stream)
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=side):
        cs.solve_update_noalloc(T, g["p"], g["rho"], Y, dt, 1.0, subDt=sub, stats=False)
    torch.cuda.synchronize()
    for r in range(3):
        T.copy_(g["T"])
        Y.copy_(g["Y"])
        sub.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(T, Te) and torch.equal(Y, Ye) and torch.equal(sub, se), r
        _ = (T * 2.0).sum().item()
    cs.set_stream(torch.cuda.current_stream().cuda_stream)


def test_noalloc_rejects():
    import torch

    from openhurricane_amd import _lib
    from openhurricane_amd._lib import OhcError
    from openhurricane_amd.chemistry import ChemistrySource, Mechanism

    mech = Mechanism.h2air()
    cs = ChemistrySource(mech, opts=_lib.default_opts(**TOL, device_pointers=1))
    small = torch.zeros(4, dtype=torch.float64, device="cuda:0")
    Ys = torch.zeros((9, 4), dtype=torch.float64, device="cuda:0")
    big = _lib.OHC_NOALLOC_MAX_CELLS + 1
    rc = _lib.lib().ohc_chem_solve_update_noalloc(
        cs._h, big, cs._p(small), cs._p(small), cs._p(small), cs._p(Ys), big, cs._p(small),
        1.0, cs._p(small), None)
    assert rc == _lib.OHC_ERR_ARG
    assert b"OHC_NOALLOC_MAX_CELLS" in _lib.lib().ohc_last_error(cs._h)
    host = ChemistrySource(mech, opts=_lib.default_opts(**TOL))
    T = np.full(4, 1000.0)
    with pytest.raises(OhcError):
        host.solve_update_noalloc(T, T.copy(), T.copy(), np.zeros((9, 4)), np.full(4, 1e-6))
