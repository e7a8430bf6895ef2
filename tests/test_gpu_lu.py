"""The integrators' LU (SquareMatrix::LUDecomposeDPivoting +
LUBacksubstituteDPivoting, SquareMatrix.hpp:185-238, LUDecompose.hpp:68-96)
through ohc_lu_solve: batched dense systems of every padded size, full and
partial (n below the padded size), well-conditioned and pivot-requiring,
against numpy.linalg.solve and the oracle's LU; a NaN column gives an all-NaN
solution."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]


@pytest.fixture(scope="module")
def cs():
    from openhurricane_amd.chemistry import ChemistrySource, Mechanism

    return ChemistrySource(Mechanism.h2air())


def _batch(A):
    """(nmat, n, n) -> the (n * n, nmat) layout of ohc_lu_solve."""
    nmat, n, _ = A.shape
    return np.ascontiguousarray(A.reshape(nmat, n * n).T)


def _systems(n, nmat, seed, kind):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((nmat, n, n))
    if kind == "dominant":
        A += n * np.eye(n)
    elif kind == "zero_diagonal":
        # every step needs a row exchange: the diagonal is zero and the
        # largest entry of each column sits off it
        A[:, np.arange(n), np.arange(n)] = 0.0
        perm = rng.permutation(n)
        A[:, np.arange(n), perm] += 4.0 * n * np.sign(rng.standard_normal((nmat, n)))
    elif kind == "graded":
        # rows scaled over 12 decades, as the Newton matrices of stiff cells
        A *= (10.0 ** rng.uniform(-6, 6, (nmat, n, 1)))
    b = rng.standard_normal((nmat, n))
    return A, b


@pytest.mark.parametrize("n", [5, 10, 12, 16, 24, 33, 40, 48, 55, 64])
@pytest.mark.parametrize("kind", ["dominant", "random", "zero_diagonal", "graded"])
def test_lu_solve_matches_numpy(cs, n, kind):
    from oracle import Oracle

    nmat = 48
    A, b = _systems(n, nmat, seed=1000 + n, kind=kind)
    print(f"lu_solve: n {n}, {kind}, {nmat} systems", flush=True)
    x = cs.lu_solve(_batch(A), np.ascontiguousarray(b.T)).T
    ref = np.linalg.solve(A, b[..., None])[..., 0]
    for m in range(nmat):
        # backward error of the partial-pivoting LU, and the forward error
        # within the condition number
        r = np.abs(A[m] @ x[m] - b[m]).max()
        assert r <= 1e-13 * (np.abs(A[m]).max() * np.abs(x[m]).max() + np.abs(b[m]).max()), (m, r)
        cond = np.linalg.cond(A[m])
        err = np.abs(x[m] - ref[m]).max() / np.abs(ref[m]).max()
        assert err <= 1e-15 * n * cond + 1e-14, (m, err, cond)
    # the oracle's restatement of the reference LU on a few of them
    for m in range(0, nmat, 12):
        xo = Oracle.lu_solve(A[m].copy(), b[m].copy())
        cond = np.linalg.cond(A[m])
        assert np.abs(x[m] - xo).max() <= (1e-15 * n * cond + 1e-14) * np.abs(xo).max()


def test_lu_solve_nan_column_gives_nan(cs):
    n, nmat = 33, 4
    A, b = _systems(n, nmat, seed=7, kind="random")
    A[0, :, 5] = np.nan    # a whole column
    A[1, 3, 3] = np.nan    # one entry: its row is never a pivot until the last column
    x = cs.lu_solve(_batch(A), np.ascontiguousarray(b.T)).T
    assert np.all(np.isnan(x[0])) and np.all(np.isnan(x[1]))
    ref = np.linalg.solve(A[2:], b[2:, :, None])[..., 0]
    assert np.allclose(x[2:], ref, rtol=1e-10, atol=0)


def test_lu_solve_arguments(cs):
    from openhurricane_amd._lib import OhcError

    A, b = _systems(4, 2, seed=3, kind="dominant")
    with pytest.raises(ValueError):
        cs.lu_solve(np.zeros((9, 2)), np.ascontiguousarray(b.T))
    with pytest.raises(OhcError):
        cs.lu_solve(np.zeros((65 * 65, 1)), np.zeros((65, 1)))
    assert cs.lu_solve(np.zeros((16, 0)), np.zeros((4, 0))).shape == (4, 0)
