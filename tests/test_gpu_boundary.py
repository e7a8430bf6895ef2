"""C-ABI boundary in host-pointer mode with species blocks whose leading
dimension exceeds the cell count (a slice of a larger cellRealArray block,
or fields with ghost cells): results equal the contiguous call and the
ld - n elements between rows are never written."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]

SENTINEL = -7.25e77


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _padded(block, ld):
    ns, n = block.shape
    big = np.full((ns, ld), SENTINEL)
    big[:, :n] = block
    return big


def test_padded_leading_dimensions_leave_padding_alone(h2mech, c1opts):
    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource
    from openhurricane_amd.sampler import h2air_states

    L = _lib.lib()
    cs = ChemistrySource(h2mech, device=0, opts=c1opts)
    st = h2air_states(h2mech.species, h2mech.tables()["W"], 96, seed=41)
    n, ns, ld = 96, h2mech.ns, 131
    dt = np.full(n, 1e-6)
    # reference: the contiguous call through the binding
    R_ref, sub_ref, _ = cs.solve(st["T"], st["p"], st["rho"], np.ascontiguousarray(st["Y"]), dt)
    Yb = _padded(st["Y"], ld)
    Rb = np.full((ns, ld), SENTINEL)
    sub = np.zeros(n)
    print(f"ohc_chem_solve n={n} ld={ld}", flush=True)
    _lib.check(L.ohc_chem_solve(cs._h, n, _p(st["T"]), _p(st["p"]), _p(st["rho"]), _p(Yb), ld,
                                _p(dt), 1.0, _p(sub), _p(Rb), ld, None), cs._h)
    assert np.all(Rb[:, n:] == SENTINEL) and np.all(Yb[:, n:] == SENTINEL)
    np.testing.assert_array_equal(Rb[:, :n], R_ref)
    np.testing.assert_array_equal(sub, sub_ref)
    # solveUpdate in place on the padded block
    T = st["T"].copy()
    sub = np.zeros(n)
    print("ohc_chem_solve_update padded", flush=True)
    _lib.check(L.ohc_chem_solve_update(cs._h, n, _p(T), _p(st["p"]), _p(st["rho"]), _p(Yb), ld,
                                       _p(dt), 1.0, _p(sub), None), cs._h)
    assert np.all(Yb[:, n:] == SENTINEL)
    T2, Y2 = st["T"].copy(), np.ascontiguousarray(st["Y"]).copy()
    cs.solve_update(T2, st["p"], st["rho"], Y2, dt)
    np.testing.assert_array_equal(Yb[:, :n], Y2)
    np.testing.assert_array_equal(T, T2)
    # source terms with the substituted diagonal (ns - 1 rows)
    Rs = np.full((ns, ld), SENTINEL)
    Ds = np.full((ns, ld), SENTINEL)
    Yb = _padded(st["Y"], ld)
    print("ohc_chem_source padded", flush=True)
    _lib.check(L.ohc_chem_source(cs._h, n, _p(st["T"]), _p(st["p"]), _p(st["rho"]), _p(Yb), ld,
                                 _p(Rs), ld, _p(Ds), ld, 1), cs._h)
    assert np.all(Rs[:, n:] == SENTINEL) and np.all(Ds[:, n:] == SENTINEL)
    assert np.all(Ds[ns - 1] == SENTINEL)  # the substituted species has no diagonal row
    R2, D2 = cs.source_terms(st["T"], st["p"], st["rho"], np.ascontiguousarray(st["Y"]),
                             diag=True, without_last=True)
    np.testing.assert_array_equal(Rs[:, :n], R2)
    np.testing.assert_array_equal(Ds[: ns - 1, :n], D2[: ns - 1])
    # mixture properties: D_k and h_k blocks
    Dk = np.full((ns, ld), SENTINEL)
    hk = np.full((ns, ld), SENTINEL)
    mu = np.zeros(n)
    print("ohc_mix_props padded", flush=True)
    _lib.check(L.ohc_mix_props(cs._h, n, _p(st["T"]), _p(st["p"]), _p(Yb), ld, _p(mu), None,
                               _p(Dk), ld, None, None, _p(hk)), cs._h)
    assert np.all(Dk[:, n:] == SENTINEL) and np.all(hk[:, n:] == SENTINEL)
    pr = cs.mix_props(st["T"], st["p"], np.ascontiguousarray(st["Y"]), kappa=False, cp=False,
                      ha=False, hk=True)
    np.testing.assert_array_equal(Dk[:, :n], pr["Dk"])
    np.testing.assert_array_equal(hk[:, :n], pr["hk"])
    np.testing.assert_array_equal(mu, pr["mu"])
