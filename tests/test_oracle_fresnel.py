"""Pins for the oracle's Fresnel propagation (Fresnel_propagate_wavefront,
WFpropagate.c:68-152, and the table variants 166-401) on the CPU."""
import numpy as np
import pytest


def _field(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2)


def _angular_spectrum(E, dx, z, lam):
    """textbook Fresnel transfer function exp(-i pi lambda z (fx^2 + fy^2))"""
    n = E.shape[0]
    f = np.fft.fftfreq(n, dx)
    H = np.exp(-1j * np.pi * lam * z * (f[None, :] ** 2 + f[:, None] ** 2))
    return np.fft.ifft2(np.fft.fft2(E) * H)


def test_z0_identity_and_unitarity(oracle):
    E = _field(64, 1)
    assert np.allclose(oracle.fresnel(E, 0.01, 0.0, 1e-6), E, atol=1e-12)
    P = oracle.fresnel(E, 0.01, 3000.0, 1e-6)
    assert np.isclose(np.linalg.norm(P), np.linalg.norm(E), rtol=1e-12)


def test_additivity_in_z(oracle):
    E = _field(64, 2)
    a = oracle.fresnel(oracle.fresnel(E, 0.02, 1200.0, 0.7e-6), 0.02, 800.0, 0.7e-6)
    b = oracle.fresnel(E, 0.02, 2000.0, 0.7e-6)
    assert np.allclose(a, b, atol=1e-11)


def test_uniform_field_is_invariant(oracle):
    E = np.full((32, 32), 0.3 - 0.4j)
    assert np.allclose(oracle.fresnel(E, 0.05, 10000.0, 2e-6), E, atol=1e-12)


@pytest.mark.parametrize("kx,ky", [(3, 0), (5, -7), (-16, 9)])
def test_plane_wave_closed_form(oracle, kx, ky):
    n, dx, z, lam = 32, 0.02, 5000.0, 1.6e-6
    y, x = np.mgrid[0:n, 0:n]
    E = np.exp(2j * np.pi * (kx * x + ky * y) / n)
    phase = -np.pi * z * lam * (kx ** 2 + ky ** 2) / (dx * n) ** 2
    assert np.allclose(oracle.fresnel(E, dx, z, lam), E * np.exp(1j * phase), atol=1e-10)


def test_matches_independent_angular_spectrum(oracle):
    E = _field(128, 3)
    for z, lam in [(1000.0, 0.5e-6), (16000.0, 2.2e-6)]:
        assert np.allclose(oracle.fresnel(E, 0.01, z, lam), _angular_spectrum(E, 0.01, z, lam), atol=1e-10)


def test_table_variant_equals_fresnel_with_exact_table(oracle):
    # Fresnel_propagate_wavefront1 with the exact (un-averaged) transfer
    # function in the FFT layout reproduces Fresnel_propagate_wavefront
    n, dx, z, lam = 64, 0.02, 3000.0, 1e-6
    E = _field(n, 4)
    k = np.fft.fftfreq(n) * n
    C = np.exp(-1j * np.pi * z * lam * (k[None, :] ** 2 + k[:, None] ** 2) / (dx * n) ** 2) / n ** 2
    assert np.allclose(oracle.fresnel1(E, C), oracle.fresnel(E, dx, z, lam), atol=1e-11)


def test_init_tf_subsampled_average(oracle):
    # WFpropagate.c:166-261: average of exp(-i coeff r^2) over a 4x4 grid of
    # sub-pixel offsets (0.5 + i - 2)/4, inside the mask radius; amplitude
    # thresholded at 0.3 to {0, 1}; divided by size^2; quadrant swapped
    size, dx, z, lam, rad = 32, 0.02, 2000.0, 1e-6, 11.0
    C = oracle.fresnel_init_tf(size, dx, z, lam, rad)
    coeff = np.pi * z * lam / (dx * size) ** 2
    off = (0.5 + np.arange(4) - 2) / 4
    exp = np.zeros((size, size), complex)
    for jj in range(size):
        for ii in range(size):
            xs = ii + off[:, None] - size // 2
            ys = jj + off[None, :] - size // 2
            r2 = xs ** 2 + ys ** 2
            s = np.sum(np.exp(-1j * coeff * r2) * (np.sqrt(r2) < rad))
            a = 1.0 if abs(s) / 16 > 0.3 else 0.0
            exp[jj, ii] = a * np.exp(1j * np.angle(s)) / size ** 2
    assert np.allclose(C, np.fft.fftshift(exp), atol=1e-14)
    # inside the mask and far from aliasing the table is the exact factor
    assert np.all(np.abs(C[0:3, 0:3]) > 0)


def test_cube(oracle):
    E = _field(32, 5)
    amp, pha = oracle.fresnel_cube(E, 0.02, 100.0, 2100.0, 4, 1e-6)
    for k, z in enumerate(100.0 + 2000.0 * np.arange(4) / 4):
        P = oracle.fresnel(E, 0.02, z, 1e-6)
        assert np.allclose(amp[k], np.abs(P), atol=1e-12)
        assert np.allclose(pha[k], np.angle(P), atol=1e-12)
