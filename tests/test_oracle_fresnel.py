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


def _tf_closed_form(size, coeff):
    """Closed form of the 4 x 4 sub-sample average of exp(-i coeff r^2) when
    the mask covers the pixel: the average separates into x and y factors,
    and pairing the offsets +-0.375 and +-0.125 gives, per axis,
        a(t) = exp(-i coeff t^2) (exp(-i coeff 0.140625) cos(0.75 coeff t)
                                  + exp(-i coeff 0.015625) cos(0.25 coeff t)) / 2
    at the pixel's signed frequency t (the quadrant swap puts t = 0 at index 0)."""
    t = np.fft.fftfreq(size) * size
    a = np.exp(-1j * coeff * t ** 2) * 0.5 * (np.exp(-1j * coeff * 0.140625) * np.cos(0.75 * coeff * t)
                                             + np.exp(-1j * coeff * 0.015625) * np.cos(0.25 * coeff * t))
    return a[None, :] * a[:, None]


@pytest.mark.parametrize("size,dx,z,lam", [(32, 0.02, 2000.0, 1e-6), (64, 0.01, 8000.0, 1.65e-6),
                                           (48, 0.005, 16000.0, 2.2e-6), (16, 0.05, 500.0, 0.5e-6)])
def test_init_tf_closed_form_without_mask(oracle, size, dx, z, lam):
    # FPMASKRAD beyond the corner of the grid: every pixel is the unit phasor
    # of the closed-form average where its modulus exceeds 0.3, else 0, / size^2
    C = oracle.fresnel_init_tf(size, dx, z, lam, 10.0 * size)
    coeff = np.pi * z * lam / (dx * size) ** 2
    A = _tf_closed_form(size, coeff)
    amp = np.abs(A)
    clear = np.abs(amp - 0.3) > 1e-9
    want = np.where(amp > 0.3, A / np.where(amp > 0, amp, 1.0), 0.0) / size ** 2
    assert np.allclose(C[clear], want[clear], atol=1e-12 / size ** 2, rtol=0)
    # the phase of a kept pixel is -coeff (t_x^2 + t_y^2) up to the sub-sample
    # spread: -coeff 2 <delta^2> = -coeff 0.15625 at small coeff t
    t = np.fft.fftfreq(size) * size
    small = (coeff * np.abs(t[None, :]) < 1e-3) & (coeff * np.abs(t[:, None]) < 1e-3)
    ph = -coeff * (t[None, :] ** 2 + t[:, None] ** 2 + 0.15625)
    assert np.allclose(np.angle(C[small] * size ** 2 * np.exp(-1j * ph[small])), 0.0, atol=1e-6)
    # the 0.3 cut-off: where the closed-form modulus is below 0.3 the table is 0
    assert np.all(C[(amp < 0.3) & clear] == 0)


def test_init_tf_aliasing_cutoff_is_reached(oracle):
    # large coeff: the sub-sample average of the chirp decays with frequency,
    # so the outer frequencies fall below 0.3 and are zeroed
    size, dx, z, lam = 64, 0.002, 20000.0, 2.2e-6
    C = oracle.fresnel_init_tf(size, dx, z, lam, 10.0 * size)
    A = _tf_closed_form(size, np.pi * z * lam / (dx * size) ** 2)
    assert np.any(np.abs(A) < 0.29) and np.any(np.abs(A) > 0.31)
    assert np.array_equal(C != 0, np.abs(A) > 0.3)


@pytest.mark.parametrize("size,rad", [(32, 7.3), (32, 11.0), (64, 20.1), (16, 3.4)])
def test_init_tf_mask_at_zero_distance(oracle, size, rad):
    # z = 0: every sub-sample adds 1, so the average is k / 16 with k the
    # sub-samples strictly inside FPMASKRAD; the 0.3 cut-off keeps a pixel
    # when k >= 5 (4 / 16 = 0.25, 5 / 16 = 0.3125), as 1 / size^2 with no phase
    C = oracle.fresnel_init_tf(size, 0.01, 0.0, 1e-6, rad)
    t = np.fft.fftfreq(size) * size
    k = np.zeros((size, size), int)
    for ox in (-0.375, -0.125, 0.125, 0.375):
        for oy in (-0.375, -0.125, 0.125, 0.375):
            k += (np.hypot(t[None, :] + ox, t[:, None] + oy) < rad)
    want = np.where(k >= 5, 1.0 / size ** 2, 0.0)
    assert np.array_equal(C, want.astype(complex))
    assert np.any((k > 0) & (k < 5))  # the ring of partly covered pixels is exercised


def test_init_tf_mask_rim_over_many_radii(oracle):
    # the strict "distance < FPMASKRAD" test of every sub-sample, over radii
    # that put sub-samples just inside and just outside the rim
    size = 32
    t = np.fft.fftfreq(size) * size
    offs = (-0.375, -0.125, 0.125, 0.375)
    d = [np.hypot(t[None, :] + ox, t[:, None] + oy) for ox in offs for oy in offs]
    radii = sorted({float(v) for v in np.concatenate([x.ravel() for x in d]) if 2 < v < 14})[::7]
    for rad in [r + e for r in radii for e in (-0.004, 0.004)]:
        C = oracle.fresnel_init_tf(size, 0.01, 0.0, 1e-6, rad)
        k = sum((x < rad).astype(int) for x in d)
        assert np.array_equal(C, np.where(k >= 5, 1.0 / size ** 2, 0.0).astype(complex)), rad


@pytest.mark.parametrize("shape", [(32, 48), (20, 20), (15, 21), (24, 9)])
def test_rectangular_and_odd_fresnel_follows_the_reference_steps(oracle, shape):
    # Fresnel_propagate_wavefront on any shape (WFpropagate.c:68-152): forward
    # FFT, permut, e^{-i coeff ((ii - n0/2)^2 + (jj - n1/2)^2)} / (n0 n1) with
    # coeff from naxes[0] on both axes, permut, inverse FFT -- built here from
    # numpy's FFT and the swap loops of permut applied to an index array
    n1, n0 = shape
    dx, z, lam = 0.02, 3000.0, 1e-6
    rng = np.random.default_rng(n0 * n1)
    E = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    def permut_np(a):
        idx = np.arange(a.size).reshape(a.shape)
        xh, yh = n0 // 2, n1 // 2
        for jj in range(yh):
            for ii in range(xh):
                idx[jj, ii], idx[jj + yh, ii + xh] = idx[jj + yh, ii + xh], idx[jj, ii]
        for jj in range(yh, n1):
            for ii in range(xh):
                idx[jj, ii], idx[jj - yh, ii + xh] = idx[jj - yh, ii + xh], idx[jj, ii]
        return a.ravel()[idx.ravel()].reshape(a.shape)

    t = permut_np(np.fft.fft2(E))
    coeff = np.pi * z * lam / (dx * n0) ** 2
    jj, ii = np.mgrid[0:n1, 0:n0]
    t = t * np.exp(-1j * coeff * ((ii - n0 // 2) ** 2 + (jj - n1 // 2) ** 2)) / (n0 * n1)
    want = np.fft.ifft2(permut_np(t)) * (n0 * n1)
    assert np.allclose(oracle.fresnel(E, dx, z, lam), want, atol=1e-10)
    assert np.allclose(oracle.permut(E), permut_np(E))
