"""Pins of the CPU reference of the surrogate network (oracle/mlp_ref.py):
its back-propagation against central finite differences of its own loss
(FP64, no BF16 rounding) and its BF16 rounding against known bit patterns."""
import numpy as np

from mlp_ref import bf16, forward_backward, gelu, gelu_grad


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 1.0e-40], dtype=np.float32)
    r = bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even (mantissa 0)
    assert r[2] == 1.0 + 2 ** -7            # above the tie -> up
    assert r[3] == -2.5
    assert np.isfinite(r[4]) or r[4] == np.inf
    assert abs(r[5]) < 1e-38


def test_gelu_derivative():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    np.testing.assert_allclose((gelu(z + h) - gelu(z - h)) / (2 * h), gelu_grad(z), rtol=1e-7, atol=1e-9)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    nin, nout, width, depth, n = 5, 4, 8, 3, 7
    npar = (nin * width + width) + (depth - 1) * (width * width + width) + (width * nout + nout)
    p = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((n, nin))
    Yt = rng.standard_normal((n, nout))
    norm = (rng.standard_normal(nin) * 0.1, rng.uniform(0.5, 2, nin), rng.standard_normal(nout) * 0.1,
            rng.uniform(0.5, 2, nout), rng.uniform(0.5, 1.5, nout))
    loss, g, _ = forward_backward(p, X, Yt, norm, nin, nout, width, depth, bf=False)
    fd = np.zeros(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = 1e-6
        lp = forward_backward(p + e, X, Yt, norm, nin, nout, width, depth, bf=False)[0]
        lm = forward_backward(p - e, X, Yt, norm, nin, nout, width, depth, bf=False)[0]
        fd[i] = (lp - lm) / 2e-6
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-8 * np.abs(fd).max())


def test_fp8_rounding_follows_the_e4m3_definition():
    """OCP E4M3: 448 is the largest finite value (saturation above), 2^-9 the
    smallest subnormal, round to nearest with ties to the even code."""
    from mlp_ref import E4M3, fp8

    assert np.nanmax(E4M3) == 448.0 and np.isnan(E4M3[0x7F]) and np.isnan(E4M3[0xFF])
    assert E4M3[1] == 2.0 ** -9 and E4M3[8] == 2.0 ** -6 and E4M3[0x38] == 1.0
    x = np.array([1.0625, 1.1875, 17.0, 100.0, -7.25, 1.5 * 2 ** -9, 0.5 * 2 ** -9, 500.0, -1e6,
                  0.3, 1e-4])
    want = np.array([1.0, 1.25, 16.0, 96.0, -7.0, 2 ** -8, 0.0, 448.0, -448.0, 0.3125, 0.0])
    np.testing.assert_array_equal(fp8(x), want)
    v = E4M3[~np.isnan(E4M3)]
    np.testing.assert_array_equal(fp8(v), v)
