"""Pins for the oracle's noise, screen synthesis (make_master_turbulence_screen,
AtmosphericTurbulence.c:643-855) and normalisation (AT.c:2246-2363)."""
import math

import numpy as np
import pytest

# Random123 known-answer vectors for Philox4x32-10
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(oracle, ctr, key, out):
    assert oracle.philox(ctr, key) == out


def test_noise_moments_and_streams(oracle):
    n = 1 << 18
    z = oracle.noise_c64(0xA75, 0, n)
    for comp in (z.real, z.imag):
        assert abs(comp.mean()) < 5 / math.sqrt(n)
        assert abs(comp.var() - 1) < 0.02
    assert abs(np.mean(z.real * z.imag)) < 5 / math.sqrt(n)
    # fourth moment of a Gaussian
    assert abs(np.mean(z.real ** 4) - 3) < 0.1
    z2 = oracle.noise_c64(0xA75, 1, n)
    assert abs(np.mean(z.real * z2.real)) < 5 / math.sqrt(n)
    assert np.array_equal(oracle.noise_c64(0xA75, 0, 1000), z[:1000])
    _, z32 = oracle.noise_c64(0xA75, 0, 1000, f32=True)
    assert np.array_equal(z32, z[:1000].astype(np.complex64))


def test_vk_filter_closed_form(oracle):
    size, L0, l0 = 64, 100.0, 2.0
    A = oracle.vk_filter(size, L0, l0)
    f0 = size / L0
    fi = (5.92 / (2 * math.pi)) * size / l0
    for ii, jj in [(0, 0), (10, 40), (33, 31), (63, 1)]:
        r2 = (ii - size // 2) ** 2 + (jj - size // 2) ** 2
        assert np.isclose(A[jj, ii], math.exp(-r2 / fi ** 2) ** 0.5 * (r2 + f0 ** 2) ** (-11 / 12), rtol=1e-13)
    assert A[size // 2, size // 2] == 0.0
    # RLIM zeroes the centre
    Ar = oracle.vk_filter(size, L0, l0, rlim=3.0)
    assert Ar[size // 2, size // 2 + 2] == 0.0 and Ar[size // 2, size // 2 + 3] > 0


def test_structure_function_direct_equals_fft_route(oracle):
    rng = np.random.default_rng(2)
    s = np.cumsum(rng.standard_normal((64, 64)), axis=1)
    D = oracle.strfn_fft(s)
    for a, b in [(1, 1), (1, 2), (2, 1), (2, 2), (5, 0), (0, 7), (63, 3)]:
        assert np.isclose(D[b, a], oracle.strfn_direct(s, a, b), rtol=1e-9)


def test_structure_function_of_a_sinusoid(oracle):
    n, k = 64, 5
    x = np.arange(n)
    s = np.tile(np.cos(2 * np.pi * k * x / n), (n, 1))
    for a in [1, 2, 3, 10]:
        assert np.isclose(oracle.strfn_direct(s, a, 0), 1 - math.cos(2 * math.pi * k * a / n), atol=1e-12)
        assert np.isclose(oracle.strfn_direct(s, 0, a), 0.0, atol=1e-14)


def test_strfn_coeff_scaling(oracle):
    rng = np.random.default_rng(3)
    s = np.cumsum(np.cumsum(rng.standard_normal((32, 32)), 0), 1)
    assert np.isclose(oracle.strfn_coeff(3.0 * s), 9.0 * oracle.strfn_coeff(s), rtol=1e-12)


def _noise(oracle, size, stream=0):
    return oracle.noise_c64(0xA75 + size, stream, size * size).reshape(size, size)


def test_master_is_normalised_and_scale_free(oracle):
    z = _noise(oracle, 128)
    o1, o2, c1, c2 = oracle.make_master(z, 1000.0, 2.0)
    # the output is divided by sqrt(C) so its own 4-lag coefficient is 1
    assert np.isclose(oracle.strfn_coeff(o1), 1.0, rtol=1e-10)
    assert np.isclose(oracle.strfn_coeff(o2), 1.0, rtol=1e-10)
    # scaling the noise scales C by the square and leaves the screen unchanged
    p1, p2, d1, d2 = oracle.make_master(2.5 * z, 1000.0, 2.0)
    assert np.allclose(p1, o1, atol=1e-10) and np.isclose(d1, 6.25 * c1, rtol=1e-10)
    # Re and Im are the two independent screens
    assert abs(np.corrcoef(o1.ravel(), o2.ravel())[0, 1]) < 0.2


def test_master_single_bin_is_a_plane_wave(oracle):
    # one noise bin at centred frequency (fx, fy): the screen is A * cos of the
    # plane wave with that frequency (the quadrant swap puts the centre at 0)
    size, fx, fy = 64, 3, 5
    z = np.zeros((size, size), complex)
    z[size // 2 + fy, size // 2 + fx] = 1.0
    o1, o2, _, _ = oracle.make_master(z, 1000.0, 2.0)
    y, x = np.mgrid[0:size, 0:size]
    wave = np.exp(-2j * np.pi * (fx * x + fy * y) / size)
    a = np.vdot(wave.real, o1) / np.vdot(wave.real, wave.real)
    assert np.allclose(o1, a * wave.real, atol=1e-10 * np.abs(a))


def test_master_structure_function_power_law(oracle):
    # large outer scale, sub-pixel inner scale: D(r) ~ r^(5/3).  An FFT screen
    # lacks power below the grid frequency, which lowers the slope slightly
    # (about 1.60 over lags 2..8 on 512^2); a wrong spectral exponent moves it
    # towards 2 (e.g. amplitude f^(-11/6)) or well below 1.5
    size = 512
    slopes = []
    for stream in range(3):
        o1, _, _, _ = oracle.make_master(_noise(oracle, size, stream), 1e6, 0.1)
        r = np.array([2, 3, 4, 6, 8])
        D = np.array([0.5 * (oracle.strfn_direct(o1, k, 0) + oracle.strfn_direct(o1, 0, k)) for k in r])
        slopes.append(np.polyfit(np.log(r), np.log(D), 1)[0])
    assert 1.5 < np.mean(slopes) < 1.72


def test_r0_measure_closed_form(oracle):
    # linear ramp g x: every 50-pixel difference is 50 g
    size, g = 128, 0.03
    s = np.tile(g * np.arange(size, dtype=float), (size, 1))
    assert np.isclose(oracle.r0_measure(s, 50), 50 * ((50 * g) ** 2 / 6.88) ** (-0.6), rtol=1e-12)
