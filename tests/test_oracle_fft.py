"""Pins for the oracle's FFT and quadrant swap (CPU only).

The reference's transforms are FFTW's unnormalised c2c with sign -1 forward
and +1 backward (src/fft/fft.c:1298-1315); the checks below fix that
convention from the DFT definition, numpy's pocketfft and closed forms."""
import numpy as np
import pytest


def _rand(shape, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 16, 64])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_matches_bruteforce_dft(oracle, n, sign):
    x = _rand(n, n)
    j = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)  # definition, evaluated by numpy
    assert np.allclose(oracle.fft1d(x, sign), W @ x, atol=1e-11 * max(n, 1))


def test_naive_dft_is_the_definition(oracle):
    x = _rand(7, 1)
    j = np.arange(7)
    assert np.allclose(oracle.dft1d_naive(x, -1), np.exp(-2j * np.pi * np.outer(j, j) / 7) @ x, atol=1e-12)


@pytest.mark.parametrize("n", [256, 1024])
def test_fft1d_matches_numpy(oracle, n):
    x = _rand(n, 3)
    assert np.allclose(oracle.fft1d(x, -1), np.fft.fft(x), atol=1e-10)
    assert np.allclose(oracle.fft1d(x, +1), np.fft.ifft(x) * n, atol=1e-10)


def test_fft2d_axes_and_sign(oracle):
    # row-major (n1 = 32 rows of n0 = 64): the 2D DFT over both axes
    x = _rand((32, 64), 4)
    assert np.allclose(oracle.fft2d(x, -1), np.fft.fft2(x), atol=1e-10)
    assert np.allclose(oracle.fft2d(x, +1), np.fft.ifft2(x) * x.size, atol=1e-10)


def test_parseval_and_roundtrip(oracle):
    x = _rand((64, 64), 5)
    X = oracle.fft2d(x, -1)
    assert np.isclose(np.sum(np.abs(X) ** 2), x.size * np.sum(np.abs(x) ** 2), rtol=1e-12)
    assert np.allclose(oracle.fft2d(X, +1) / x.size, x, atol=1e-12)


def test_impulse_constant_and_plane_wave(oracle):
    n = 32
    d = np.zeros((n, n), complex)
    d[0, 0] = 1.0
    assert np.allclose(oracle.fft2d(d, -1), np.ones((n, n)))
    assert np.allclose(oracle.fft2d(np.ones((n, n), complex), -1), n * n * d)
    # exp(+2 pi i (kx x + ky y)/n) has a single forward bin at (ky, kx)
    kx, ky = 5, 27
    y, x = np.mgrid[0:n, 0:n]
    pw = np.exp(2j * np.pi * (kx * x + ky * y) / n)
    X = oracle.fft2d(pw, -1)
    expect = np.zeros((n, n), complex)
    expect[ky, kx] = n * n
    assert np.allclose(X, expect, atol=1e-9)


def _permut_loops(a):
    """the two swap loops of fft.c:698-711, written out in Python"""
    a = a.copy()
    n1, n0 = a.shape
    xh, yh = n0 // 2, n1 // 2
    for jj in range(yh):
        for ii in range(xh):
            a[jj, ii], a[jj + yh, ii + xh] = a[jj + yh, ii + xh], a[jj, ii]
    for jj in range(yh, n1):
        for ii in range(xh):
            a[jj, ii], a[jj - yh, ii + xh] = a[jj - yh, ii + xh], a[jj, ii]
    return a


@pytest.mark.parametrize("shape", [(8, 8), (6, 10), (16, 4)])
def test_permut_even_is_fftshift_involution(oracle, shape):
    x = _rand(shape, 6)
    p = oracle.permut(x)
    assert np.array_equal(p, np.fft.fftshift(x))
    assert np.array_equal(oracle.permut(p), x)
    r = x.real.copy()
    assert np.array_equal(oracle.permut(r), np.fft.fftshift(r))


@pytest.mark.parametrize("shape", [(5, 5), (7, 4), (4, 9)])
def test_permut_odd_follows_swap_loops(oracle, shape):
    x = _rand(shape, 7)
    assert np.array_equal(oracle.permut(x), _permut_loops(x))


@pytest.mark.parametrize("shape", [(12, 20), (9, 15), (64, 48), (1, 30), (17, 1)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft2d_rectangular_and_odd_matches_numpy(oracle, shape, sign):
    # FFTW's 2D plan over any (n1, n0) shape (fft.c:1190): numpy's pocketfft
    # as the independent FP64 check
    rng = np.random.default_rng(sum(shape) + sign)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    want = np.fft.fft2(x) if sign < 0 else np.fft.ifft2(x) * x.size
    assert np.allclose(oracle.fft2d(x, sign), want, atol=1e-9)
