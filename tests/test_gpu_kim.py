"""The GPU path's source terms with the "Kim" diagonal (ohc_chem_source_kim:
CUDAChemistrySource::calculateSourceTermsImp[Async][Hybrid],
CUDAChemSourceKimDiagonalKernel.cu:39-235) against the oracle: R and the
FP64 diagonal to 1e-9 of each cell's largest entry, the FP32 (Hybrid)
diagonal against the FP64 oracle to 1e-5 (its stated error), on H2-air and
the 32- and 54-species synthetic mechanisms, with a padded leading dimension.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]


def _case(kind, n):
    from oracle import Oracle

    from openhurricane_amd.chemistry import Mechanism
    from openhurricane_amd.sampler import h2air_states, synthetic_workload

    if kind == "h2air":
        mech = Mechanism.h2air()
        st = h2air_states(mech.species, mech.tables()["W"], n, seed=51)
    else:
        ns, nr = (32, 150) if kind == "syn32" else (54, 270)
        mech = Mechanism.synthetic(ns, nr, seed=1 if ns == 32 else 2)
        st = synthetic_workload(mech.species, mech.tables()["W"], n, seed=52, front=0.3)
    return mech, Oracle(mech), st


def _rel(a, b):
    return np.max(np.abs(a - b) / (np.abs(b).max(axis=0) + 1e-300))


@pytest.mark.parametrize("kind", ["h2air", "syn32", "syn54"])
def test_kim_source_matches_oracle(kind):
    from openhurricane_amd.chemistry import ChemistrySource

    n = 1000
    mech, orc, st = _case(kind, n)
    cs = ChemistrySource(mech)
    print(f"kim {kind}: {n} cells, integrator kernel {cs.kernel_name}", flush=True)
    Ro, Do = orc.chem_source_kim(st["T"], st["p"], st["rho"], st["Y"])
    R, D = cs.source_terms_kim(st["T"], st["p"], st["rho"], st["Y"])
    eR, eD = _rel(R, Ro), _rel(D, Do)
    R32, D32 = cs.source_terms_kim(st["T"], st["p"], st["rho"], st["Y"], fp32=True)
    e32 = _rel(D32.astype(np.float64), Do)
    print(f"max rel err: R {eR:.2e}, diag FP64 {eD:.2e}, diag FP32 {e32:.2e}")
    assert eR <= 1e-9 and eD <= 1e-9
    assert D32.dtype == np.float32 and e32 <= 1e-5
    np.testing.assert_array_equal(R32, R)


def test_kim_padded_leading_dimension():
    """Species blocks with ld > n: the padding of every row is left as it
    was (host-pointer staging copies n entries per row)."""
    import ctypes as C

    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import ChemistrySource

    n, ld = 300, 317
    mech, orc, st = _case("syn32", n)
    cs = ChemistrySource(mech)
    ns = mech.ns
    Y = np.full((ns, ld), 7.0)
    Y[:, :n] = st["Y"]
    R = np.full((ns, ld), -3.0)
    D = np.full((ns - 1, ld), -5.0, dtype=np.float32)
    rc = _lib.lib().ohc_chem_source_kim(cs._h, n, cs._p(st["T"]), cs._p(st["p"]),
                                        cs._p(st["rho"]), C.c_void_p(Y.ctypes.data), ld,
                                        C.c_void_p(R.ctypes.data), ld,
                                        C.c_void_p(D.ctypes.data), ld, 1)
    _lib.check(rc, cs._h)
    Ro, Do = orc.chem_source_kim(st["T"], st["p"], st["rho"], st["Y"])
    assert _rel(R[:, :n], Ro) <= 1e-9 and _rel(D[:, :n].astype(np.float64), Do) <= 1e-5
    assert np.all(R[:, n:] == -3.0) and np.all(D[:, n:] == -5.0)
