"""Pins for the oracle's layer set-up, motion, sampling, frame loop and the
two unwrap rules (AtmosphericTurbulence.c:1947-4262)."""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("alt,cn2,bin_m,sa,sli", [
    ([0, 5000, 10000], [1, 1, 1], 1000.0, [0, 5000, 10000], [0, 1, 2]),
    ([0, 500, 5000], [1, 1, 1], 1000.0, [250, 5000], [0, 0, 1]),
    # 0 and 100 merge first (weight 1+1 at 50), then 50 and 200 (2+2 at 125)
    ([0, 100, 200, 5000], [1, 1, 2, 1], 1000.0, [125, 5000], [0, 0, 0, 1]),
    ([1000], [1], 1000.0, [1000], [0]),
    ([0, 800, 1600, 2400], [1, 1, 1, 1], 500.0, [0, 800, 1600, 2400], [0, 1, 2, 3]),
])
def test_superlayers_hand_worked(oracle, alt, cn2, bin_m, sa, sli):
    got_sa, got_sli = oracle.superlayers(alt, cn2, bin_m)
    assert np.allclose(got_sa, sa)
    assert list(got_sli) == sli


def test_advance_position_wraps(oracle):
    M = 64
    assert oracle.advance_position(63.5, 10.0, 100.0, 0.0, 0.01, M) == pytest.approx((0.5, 10.0))
    assert oracle.advance_position(0.25, 1.0, -50.0, -200.0, 0.01, M) == pytest.approx((63.75, 63.0))
    # x == M is left alone (the reference wraps only x > M)
    assert oracle.advance_position(63.0, 0.0, 100.0, 0.0, 0.01, M)[0] == pytest.approx(64.0)


@pytest.mark.parametrize("bicubic", [True, False])
def test_sampler_integer_shift_is_a_copy(oracle, bicubic):
    M = 32
    F = np.random.default_rng(0).standard_normal((M, M))
    for (x, y), (ii, jj) in [((3.0, 7.0), (0, 0)), ((30.0, 2.0), (5, 9)), ((0.0, 31.0), (1, 4))]:
        got = oracle.sample(F, x, y, ii, jj, bicubic)
        assert got == pytest.approx(F[(int(y) + jj) % M, (int(x) + ii) % M], abs=1e-12)


def test_bilinear_exact_on_affine(oracle):
    M = 32
    y, x = np.mgrid[0:M, 0:M].astype(float)
    F = 0.3 + 1.7 * x - 0.45 * y
    for xp, yp in [(2.25, 3.5), (10.8, 0.1), (5.0, 7.75)]:
        for ii, jj in [(0, 0), (4, 9)]:
            X, Y = xp + ii, yp + jj
            assert oracle.sample(F, xp, yp, ii, jj, False) == pytest.approx(0.3 + 1.7 * X - 0.45 * Y, abs=1e-12)


def test_catmull_rom_exact_on_quadratics(oracle):
    M = 32
    y, x = np.mgrid[0:M, 0:M].astype(float)
    for fn in [lambda X, Y: X * X, lambda X, Y: X * Y, lambda X, Y: Y * Y - 2 * X + 1.5]:
        F = fn(x, y)
        for xp, yp in [(3.3, 4.6), (8.75, 2.2)]:
            assert oracle.sample(F, xp, yp, 3, 5, True) == pytest.approx(fn(xp + 3, yp + 5), abs=1e-10)


def test_catmull_rom_is_not_exact_on_cubics(oracle):
    # Catmull-Rom reproduces quadratics only; at t = 1/4 a cubic shows its
    # error, which pins the exact weights (a cubic B-spline or Lagrange
    # weights give other values)
    M = 32
    x = np.tile(np.arange(M, dtype=float), (M, 1))
    F = x ** 3
    X = 5.25
    got = oracle.sample(F, X, 2.0, 0, 0, True)
    # taps 4..7 with weights (-9, 111, 29, -3) / 128 at t = 1/4
    assert got == pytest.approx((-9 * 64 + 111 * 125 + 29 * 216 - 3 * 343) / 128, abs=1e-12)
    assert abs(got - X ** 3) > 0.01


@pytest.mark.parametrize("bicubic", [True, False])
def test_sampler_periodic_wrap(oracle, bicubic):
    # a periodic field sampled across the master edge equals its continuation
    M = 32
    y, x = np.mgrid[0:M, 0:M].astype(float)
    k = 2 * np.pi / M
    F = np.cos(k * x) + 0.5 * np.sin(2 * k * y)
    G = np.tile(F, (2, 2))
    for xp, yp, ii, jj in [(30.4, 29.7, 1, 2), (31.9, 0.3, 0, 0), (15.5, 31.2, 16, 0)]:
        a = oracle.sample(F, xp, yp, ii, jj, bicubic)
        b = oracle.sample(G, xp, yp, ii, jj, bicubic)
        assert a == pytest.approx(b, abs=1e-12)


def _sim(oracle, masters, fresnel=True, amplitude=True, unwrap=(0,), bicubic=True, alts=None):
    L = masters.shape[0]
    alts = np.arange(L) * 3000.0 if alts is None else np.asarray(alts, float)
    sa, sli = oracle.superlayers(alts, np.ones(L), 1000.0)
    W = len(unwrap)
    return oracle.Sim(masters.shape[1] // 2, masters.shape[1] // 2, masters, sa, sli, [0.5e-6] * W,
                      np.ones((W, L)), list(unwrap), np.full(L, 1.6), np.full(L, 2.3),
                      np.linspace(10, 30, L), np.linspace(-5, 8, L), 0.02, 0.01, 0.0, bicubic,
                      fresnel=fresnel, amplitude=amplitude)


def _masters(L, M, seed=0, amp=6.0):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:M, 0:M]
    out = []
    for _ in range(L):
        acc = np.zeros((M, M))
        for _ in range(6):
            kx, ky = rng.integers(-3, 4, 2)
            acc += rng.standard_normal() * np.cos(2 * np.pi * (kx * x + ky * y) / M + rng.uniform(0, 6.3))
        out.append(amp * acc)
    return np.stack(out)


def test_single_layer_field_is_exp_i_screen(oracle):
    m = _masters(1, 32)
    s = _sim(oracle, m)
    E, PHI, pha, amp = s.frame()
    assert np.allclose(E[0], np.exp(1j * PHI[0]), atol=1e-12)
    # PHI is the sampled screen at the advanced position
    v = np.array([[oracle.sample(m[0], s.xpos[0], s.ypos[0], ii, jj, True) for ii in range(16)] for jj in range(16)])
    assert np.allclose(PHI[0], v, atol=1e-12)
    # with E = e^{i PHI} the reference-lambda unwrap returns PHI itself
    assert np.allclose(pha[0], PHI[0], atol=1e-9)
    assert np.allclose(amp[0], 1.0, atol=1e-12)


def test_propagation_conserves_energy_and_phase_accumulates(oracle):
    m = _masters(3, 64, seed=1)
    s = _sim(oracle, m)
    E, PHI, _, _ = s.frame()
    assert np.isclose(np.sum(np.abs(E[0]) ** 2), E[0].size, rtol=1e-10)
    v = sum(np.array([[oracle.sample(m[l], s.xpos[l], s.ypos[l], ii, jj, True) for ii in range(32)]
                      for jj in range(32)]) for l in range(3))
    assert np.allclose(PHI[0], v, atol=1e-10)


def test_layer_loop_matches_numpy_restatement(oracle):
    # the top super-layer's screen, Fresnel to the next, its screen, ... with
    # numpy's FFT: an independent implementation of the loop structure
    m = _masters(3, 64, seed=2)
    s = _sim(oracle, m)
    E, _, _, _ = s.frame()
    n = 32
    f = np.fft.fftfreq(n, 0.02)
    E2 = np.ones((n, n), complex)
    for l in (2, 1, 0):
        if l != 2:
            H = np.exp(-1j * np.pi * 0.5e-6 * 3000.0 * (f[None, :] ** 2 + f[:, None] ** 2))
            E2 = np.fft.ifft2(np.fft.fft2(E2) * H)
        v = np.array([[oracle.sample(m[l], s.xpos[l], s.ypos[l], ii, jj, True) for ii in range(n)]
                      for jj in range(n)])
        E2 = E2 * np.exp(1j * v)
    assert np.allclose(E[0], E2, atol=1e-10)


def test_no_fresnel_and_phase_only_modes(oracle):
    m = _masters(3, 32, seed=3)
    E, PHI, pha, amp = _sim(oracle, m, fresnel=False).frame()
    assert np.allclose(E[0], np.exp(1j * PHI[0]), atol=1e-12)
    _, PHI2, pha2, amp2 = _sim(oracle, m, amplitude=False).frame()
    assert np.allclose(pha2[0], PHI2[0])
    assert np.all(amp2 == 1.0)


def test_simple_unwrap_rule(oracle):
    # pha = atan2 + 2 pi ((long)(PHI/2pi + 1000.5) - 1000): PHI's nearest
    # multiple of 2 pi, not the one nearest to PHI - atan2
    n = 8
    PHI = np.full((n, n), 3.0)
    E = np.exp(1j * np.full((n, n), 3.4))  # atan2 = 3.4 - 2 pi
    pha, amp = oracle.extract_simple(E, PHI, n)
    assert np.allclose(pha, 3.4 - 2 * np.pi)
    PHI = np.full((n, n), 4 * np.pi + 0.2)
    E = np.exp(1j * PHI) * 0.5
    pha, amp = oracle.extract_simple(E, PHI, n)
    assert np.allclose(pha, PHI) and np.allclose(amp, 0.5)


def test_block_unwrap_identity_on_consistent_fields(oracle):
    n = 64
    y, x = np.mgrid[0:n, 0:n]
    PHI = 25.0 * np.sin(x / 9.0) + 0.2 * y
    pha, amp, peak, it = oracle.extract_block(np.exp(1j * PHI), PHI, n)
    assert np.allclose(pha, PHI, atol=1e-9)
    assert np.allclose(peak, 0.0, atol=1e-9)


def test_block_unwrap_heuristic_fixes_an_isolated_block(oracle):
    # field phase = PHI + 0.9 pi everywhere except one 20 x 20 block at
    # PHI + 1.1 pi: that block's peak comes out at +0.9 pi against -0.9 pi
    # around it; the +-2 pi fix-up flips it, so the output is PHI + offset
    # with no 2 pi jump
    n = 200
    y, x = np.mgrid[0:n, 0:n]
    PHI = 12.0 * np.cos(y / 30.0) + 0.05 * x
    offset = np.full((n, n), 0.9 * np.pi)
    offset[100:120, 100:120] = 1.1 * np.pi
    pha, amp, peak, it = oracle.extract_block(np.exp(1j * (PHI + offset)), PHI, n)
    assert it == 2
    assert np.isclose(peak[5, 5], -1.1 * np.pi) and np.isclose(peak[4, 4], -0.9 * np.pi)
    assert np.allclose(pha, PHI + offset, atol=1e-9)
    # without the fix-up the block would be unwrapped onto the other branch
    assert np.max(np.abs(np.diff(pha, axis=1))) < np.pi


def test_block_unwrap_straight_branch_cut_does_not_converge(oracle):
    # a straight cut through pi makes whole columns of blocks flip back and
    # forth; the reference loop stops at its 1000-iteration cap
    n = 200
    y, x = np.mgrid[0:n, 0:n]
    PHI = 12.0 * np.cos(y / 30.0)
    offset = 0.6 * np.pi + 0.8 * np.pi * x / n
    pha, amp, peak, it = oracle.extract_block(np.exp(1j * (PHI + offset)), PHI, n)
    assert it == 1000


def test_block_unwrap_small_output_has_no_blocks(oracle):
    n = 16  # Nout / 20 = 0 blocks: peak stays 0
    PHI = np.linspace(0, 30, n * n).reshape(n, n)
    pha, _, peak, it = oracle.extract_block(np.exp(1j * PHI), PHI, n)
    assert peak.size == 0 and it == 1
    assert np.allclose(pha, PHI, atol=1e-9)
