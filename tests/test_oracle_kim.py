"""Pins of the oracle's restatement of the GPU path's source terms with the
"Kim" diagonal (CUDAChemistrySource::calculateSourceTermsImp[Async][Hybrid],
CUDAChemistrySource.cpp:152-250; calcChemSourceTermsKimDiagonal,
CUDAChemSourceKimDiagonalKernel.cu:39-105; reactionTable::qfr<dOmegaType>,
reactionTableCUDA.inl:142-273).

* On a mechanism without third-body reactions the diagonal is the
  mass-action d(omega_k)/d(c_k) with the last species eliminated through
  c_ns = (rho - sum_j W_j c_j) / W_ns, i.e. J_kk - (W_k / W_ns) J_k,ns of the
  analytic Jacobian at fixed T (finite-difference pinned in
  test_oracle_pins).
* R_k = omega_k W_k is the CPU path's source term.
* The FP32 ("Hybrid") diagonal is the FP64 one to float rounding.
"""
import numpy as np


def _cells(mech, n, seed):
    from openhurricane_amd.sampler import synthetic_workload

    st = synthetic_workload(mech.species, mech.tables()["W"], n, seed=seed, front=0.5)
    return st


def test_kim_diagonal_is_substituted_jacobian_diagonal():
    from oracle import Oracle

    from openhurricane_amd import _lib
    from openhurricane_amd.chemistry import Mechanism

    mech = Mechanism.synthetic(32, 150, seed=7, frac_rev=0.2, frac_tb=0.0)
    orc = Oracle(mech)
    W = mech.tables()["W"]
    n, ns = 8, 32
    st = _cells(mech, n, 9)
    R, D = orc.chem_source_kim(st["T"], st["p"], st["rho"], st["Y"])
    Rc = orc.chem_source(_lib.default_opts(), st["T"], st["p"], st["rho"], st["Y"])
    np.testing.assert_array_equal(R, Rc)
    for i in range(n):
        c = st["rho"][i] * st["Y"][:, i] / W
        _, J = orc.jacobian(np.concatenate([c, [st["T"][i]]]), st["p"][i])
        a = np.diag(J)[: ns - 1]
        b = (W[: ns - 1] / W[ns - 1]) * J[: ns - 1, ns - 1]
        want = a - b
        scale = np.maximum(np.abs(a), np.abs(b)) + 1e-300
        assert np.max(np.abs(D[:, i] - want) / scale) <= 1e-12


def test_kim_hybrid_is_fp64_to_float_rounding(h2mech, h2oracle):
    from oracle import Oracle

    from openhurricane_amd.chemistry import Mechanism
    from openhurricane_amd.sampler import h2air_states

    st = h2air_states(h2mech.species, h2mech.tables()["W"], 64, seed=4)
    for orc, s in ((h2oracle, st), (None, None)):
        if orc is None:
            mech = Mechanism.synthetic(32, 150, seed=1)
            orc, s = Oracle(mech), _cells(mech, 64, 10)
        R64, D64 = orc.chem_source_kim(s["T"], s["p"], s["rho"], s["Y"])
        R32, D32 = orc.chem_source_kim(s["T"], s["p"], s["rho"], s["Y"], fp32=True)
        assert D32.dtype == np.float32
        np.testing.assert_array_equal(R32, R64)
        scale = np.abs(D64).max(axis=0) + 1e-300
        err = np.abs(D32.astype(np.float64) - D64) / scale
        assert err.max() <= 1e-5, err.max()
        assert err.max() > 0.0  # the float path is really float
