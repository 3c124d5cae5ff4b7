"""The residue decompositions behind the pruned transforms, checked against numpy's FFT.

The column pass and the adjoint passes of the real, even profile spectrum split an N-point
transform (N = 512 P) by residue classes k0 (mod P).  With W_N = exp(-2 pi i / N):

* forward (column pass): an even sequence x supported on |n| < 256 has
  X[P K + k0] = sum_{m=0}^{511} y_k0[m] W_512^(m K), where y_k0[m] = x[m'] W_N^(m' k0) and m' is m
  (m < 256) or m - 512: a 512-point transform at the fractional frequencies K + k0/P.  Evenness
  gives X[N - k] = X[k], so residues 0..P/2 cover every output.
* adjoint (adjoint column and row passes): for an even sequence z of length N, with Y_k0 the
  512-point transform of z[k0 + P K], the outputs m = 0..256 are
  Z[m] = sum_{k0=0}^{P/2} W_N^(m k0) Y_k0[m] + [k0 != 0, P/2] W_N^(-m k0) Y_k0[512 - m]
  (the second term standing for residue P - k0, whose sequence is residue k0's reversed).

Both are restated here in numpy, as the kernels compute them, and compared with numpy.fft on
random even sequences; a sign, index or mirror mistake in either fails them.
"""
import numpy as np
import pytest


def _even_window(rng, N, rs):
    """An even length-N sequence (x[N - n] = x[n]) supported on |n| < rs."""
    x = np.zeros(N, dtype=complex)
    half = rng.standard_normal(rs) + 1j * rng.standard_normal(rs)
    x[:rs] = half
    x[N - rs + 1:] = half[1:][::-1]
    return x


def _forward_by_residues(x, N):
    P = N // 512
    mp = np.where(np.arange(512) < 256, np.arange(512), np.arange(512) - 512)   # signed position m'
    X = np.zeros(N, dtype=complex)
    for k0 in range(P // 2 + 1):
        y = x[mp % N] * np.exp(-2j * np.pi * mp * k0 / N)
        Yk = np.fft.fft(y)                                   # K = 0..511
        k = P * np.arange(512) + k0
        X[k] = Yk
        X[(N - k) % N] = Yk                                  # evenness: residue P - k0
    return X


def _adjoint_by_residues(z, N, mmax=256):
    P = N // 512
    m = np.arange(mmax + 1)
    Z = np.zeros(mmax + 1, dtype=complex)
    for k0 in range(P // 2 + 1):
        Y = np.fft.fft(z[k0 + P * np.arange(512)])
        Z += np.exp(-2j * np.pi * m * k0 / N) * Y[m % 512]
        if k0 != 0 and 2 * k0 != P:
            Z += np.exp(2j * np.pi * m * k0 / N) * Y[(512 - m) % 512]
    return Z


@pytest.mark.parametrize("N", [1024, 2048, 4096, 8192])
@pytest.mark.parametrize("rs", [1, 37, 201, 256])
def test_forward_residue_transform(N, rs):
    rng = np.random.default_rng(N + rs)
    x = _even_window(rng, N, rs)
    ref = np.fft.fft(x)
    got = _forward_by_residues(x, N)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("N", [1024, 2048, 4096, 8192])
def test_adjoint_residue_transform(N):
    rng = np.random.default_rng(N)
    z = _even_window(rng, N, N // 2 + 1)                     # a full even sequence
    ref = np.fft.fft(z)[:257]
    got = _adjoint_by_residues(z, N)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_mirror_term_is_needed():
    """Dropping the mirror term (residues P - k0) breaks the adjoint identity."""
    N = 4096
    P = N // 512
    z = _even_window(np.random.default_rng(7), N, N // 2 + 1)
    ref = np.fft.fft(z)[:257]
    m = np.arange(257)
    Z = np.zeros(257, dtype=complex)
    for k0 in range(P // 2 + 1):
        Z += np.exp(-2j * np.pi * m * k0 / N) * np.fft.fft(z[k0 + P * np.arange(512)])[m % 512]
    assert np.max(np.abs(Z - ref)) > 1e-3 * np.max(np.abs(ref))
