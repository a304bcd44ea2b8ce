"""GPU parity of the transforms (FFT, permut, Fresnel, noise, screen synthesis)
against the CPU oracle, through the C ABI (libatsim.so).

Tolerances: FP32 butterflies vs the FP64 oracle; relative L2 <= 2e-6 for a
2D FFT, <= 1e-5 for Fresnel steps and screens (the bar for FP32 outputs)."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a, dtype=np.complex64):
    return torch.from_numpy(np.ascontiguousarray(a, dtype)).cuda()


def _field(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512, 1024, 2048])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft2d_vs_oracle(ctx, oracle, n, sign):
    x = _field(n, n + (sign > 0))
    ref = oracle.fft2d(x.astype(np.complex128), sign)
    t = _dev(x)
    ctx.fft2d(t, sign)
    assert rel_l2(t.cpu().numpy(), ref) < 2e-6


@pytest.mark.parametrize("n", [4096, 8192])
def test_fft2d_large_vs_numpy(ctx, n):
    x = _field(n, 7)
    t = _dev(x)
    ctx.fft2d(t, -1)
    ref = np.fft.fft2(x.astype(np.complex128))
    assert rel_l2(t.cpu().numpy(), ref) < 3e-6


def test_fft2d_batched(ctx, oracle):
    xs = np.stack([_field(64, s) for s in range(5)])
    t = _dev(xs)
    ctx.fft2d(t, -1)
    got = t.cpu().numpy()
    for k in range(5):
        assert rel_l2(got[k], oracle.fft2d(xs[k].astype(np.complex128), -1)) < 2e-6


def test_fft_roundtrip(ctx):
    x = _field(256, 3)
    t = _dev(x)
    ctx.fft2d(t, -1)
    ctx.fft2d(t, +1)
    assert rel_l2(t.cpu().numpy() / 256 ** 2, x) < 2e-6


@pytest.mark.parametrize("shape", [(64, 64), (32, 128), (256, 256)])
def test_permut_vs_oracle(ctx, oracle, shape):
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    t = _dev(x)
    ctx.permut(t)
    assert np.array_equal(t.cpu().numpy(), oracle.permut(x.astype(np.complex128)).astype(np.complex64))


@pytest.mark.parametrize("n,z,lam", [(64, 1000.0, 0.5e-6), (256, 5000.0, 1.65e-6), (1024, 2222.0, 2.2e-6),
                                     (2048, 2222.0, 0.5e-6)])
def test_fresnel_vs_oracle(ctx, oracle, n, z, lam):
    rng = np.random.default_rng(n)
    ph = np.cumsum(np.cumsum(rng.standard_normal((n, n)) * 0.05, 0), 1)
    E = (np.exp(1j * ph) * (1 + 0.1 * rng.standard_normal((n, n)))).astype(np.complex64)
    dx = 0.01
    ref = oracle.fresnel(E.astype(np.complex128), dx, z, lam)
    got = ctx.fresnel(_dev(E), dx, z, lam).cpu().numpy()
    err = rel_l2(got, ref)
    assert err < 1e-5, err


def test_fresnel_batch_and_inplace(ctx, oracle):
    n = 128
    Es = np.stack([_field(n, s) for s in range(3)])
    t = _dev(Es)
    ctx.fresnel(t, 0.02, 3000.0, 1e-6, out=t)
    got = t.cpu().numpy()
    for k in range(3):
        assert rel_l2(got[k], oracle.fresnel(Es[k].astype(np.complex128), 0.02, 3000.0, 1e-6)) < 1e-5


def test_fresnel_tf_init_apply_cube(ctx, oracle):
    n, dx, z, lam, rad = 128, 0.02, 4000.0, 0.8e-6, 50.0
    Cg = ctx.fresnel_tf_init(n, dx, z, lam, rad).cpu().numpy()
    Co = oracle.fresnel_init_tf(n, dx, z, lam, rad)
    assert rel_l2(Cg, Co) < 1e-6
    E = _field(n, 11)
    got = ctx.fresnel_tf_apply(_dev(E), _dev(Co)).cpu().numpy()
    assert rel_l2(got, oracle.fresnel1(E.astype(np.complex128), Co)) < 1e-5
    amp, pha = ctx.fresnel_cube(_dev(E), dx, 100.0, 4000.0, 3, lam)
    ao, po = oracle.fresnel_cube(E.astype(np.complex128), dx, 100.0, 4000.0, 3, lam)
    assert rel_l2(amp.cpu().numpy(), ao) < 1e-5
    # phase compared as unit phasors (branch cut at +-pi)
    assert rel_l2(np.exp(1j * pha.cpu().numpy()), np.exp(1j * po)) < 1e-5


def test_noise_matches_oracle(ctx, oracle):
    n = 1 << 16
    g = ctx.noise(n, 0xA75, 3).cpu().numpy()
    _, o32 = oracle.noise_c64(0xA75, 3, n, f32=True)
    assert np.max(np.abs(g - o32)) < 2e-6
    assert np.mean(g == o32) > 0.999


@pytest.mark.parametrize("size,L0,l0", [(256, 1000.0, 1.0), (1024, 2000.0, 0.5), (4096, 2000.0, 1.0),
                                        (8192, 2000.0, 1.0)])
def test_make_master_vs_oracle(ctx, oracle, size, L0, l0):
    # 8192^2 is the master size of the 2048^2 headline workload (L0 = 20 m and
    # l0 = 1 cm at 1 cm pixels)
    print(f"make_master {size}^2 L0 {L0} px l0 {l0} px", flush=True)
    z = ctx.noise(size * size, 0xA75 + size, 0).reshape(size, size)
    zn = z.cpu().numpy().astype(np.complex128)
    o1, o2, c1, c2 = oracle.make_master(zn, L0, l0)
    g1, g2, d1, d2 = ctx.make_master(z, L0, l0)
    assert abs(d1 / c1 - 1) < 1e-5 and abs(d2 / c2 - 1) < 1e-5
    assert rel_l2(g1.cpu().numpy(), o1) < 1e-5
    assert rel_l2(g2.cpu().numpy(), o2) < 1e-5


def test_r0_and_strfn_vs_oracle(ctx, oracle):
    rng = np.random.default_rng(5)
    s = np.cumsum(np.cumsum(rng.standard_normal((256, 256)), 0), 1).astype(np.float32)
    t = _dev(s, np.float32)
    assert abs(ctx.r0_measure(t, 50) / oracle.r0_measure(s.astype(np.float64), 50) - 1) < 1e-9
    for a, b in [(1, 1), (2, 1), (0, 3)]:
        assert abs(ctx.strfn(t, a, b) / oracle.strfn_direct(s.astype(np.float64), a, b) - 1) < 1e-9
