"""Transforms of any image shape on the GPU (atsim_fft2d_nm, atsim_permut on
odd sizes, atsim_fresnel_nm) against the oracle, which plans any shape as
FFTW does (fft.c:1145-1209) and follows WFpropagate.c:68-152 on any shape.

Bars: relative L2 <= 1e-5 for transforms and propagations (FP32 arithmetic,
Bluestein chirps from FP64 tables); permut bit-exact."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("shape", [(160, 96), (100, 100), (384, 384), (256, 128), (128, 243), (97, 61), (24, 1000),
                                   (1024, 1), (3, 5)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft2d_any_shape_vs_oracle(ctx, oracle, shape, sign):
    print(f"fft2d_nm shape (n1, n0) = {shape} sign {sign}", flush=True)
    x = _rand(shape, sum(shape) + sign)
    g = ctx.fft2d_nm(torch.tensor(x, device="cuda"), sign).cpu().numpy()
    want = oracle.fft2d(x.astype(np.complex128), sign) if shape[0] * shape[1] <= 40000 else \
        (np.fft.fft2(x.astype(np.complex128)) if sign < 0 else np.fft.ifft2(x.astype(np.complex128)) * x.size)
    e = rel_l2(g, want)
    print("  rel L2", e, flush=True)
    assert e < 1e-5


def test_fft2d_any_shape_batched_and_large(ctx):
    # a batch of 3 rectangular images, and 1000^2 / 4096 x 3000 against
    # numpy (pocketfft, FP64): Bluestein at lengths 1000, 3000 (M = 2048, 8192)
    x = _rand((3, 72, 200), 1)
    g = ctx.fft2d_nm(torch.tensor(x, device="cuda"), -1).cpu().numpy()
    assert rel_l2(g, np.fft.fft2(x.astype(np.complex128))) < 1e-5
    for shape in [(1000, 1000), (3000, 4096)]:
        print(f"fft2d_nm large {shape}", flush=True)
        x = _rand(shape, 2)
        g = ctx.fft2d_nm(torch.tensor(x, device="cuda"), -1).cpu().numpy()
        e = rel_l2(g, np.fft.fft2(x.astype(np.complex128)))
        print("  rel L2", e, flush=True)
        assert e < 1e-5
    # lengths past the Bluestein limit are refused
    from atmosphericturbulencesimul_amd._lib import AtsimError
    with pytest.raises(AtsimError):
        ctx.fft2d_nm(torch.zeros((4, 5000), dtype=torch.complex64, device="cuda"), -1)


@pytest.mark.parametrize("shape", [(61, 97), (5, 5), (9, 4), (64, 33)])
def test_permut_odd_sizes_vs_oracle(ctx, oracle, shape):
    x = _rand(shape, 3)
    g = ctx.permut(torch.tensor(x, device="cuda")).cpu().numpy()
    assert np.array_equal(g, oracle.permut(x.astype(np.complex128)).astype(np.complex64))


@pytest.mark.parametrize("shape,z,lam", [((160, 96), 2000.0, 1e-6), ((61, 97), 5000.0, 1.65e-6),
                                         ((384, 384), 1000.0, 0.5e-6), ((128, 256), 3000.0, 2.2e-6)])
def test_fresnel_any_shape_vs_oracle(ctx, oracle, shape, z, lam):
    print(f"fresnel_nm shape {shape}", flush=True)
    x = _rand(shape, 4)
    g = ctx.fresnel_nm(torch.tensor(x, device="cuda"), 0.02, z, lam).cpu().numpy()
    e = rel_l2(g, oracle.fresnel(x.astype(np.complex128), 0.02, z, lam))
    print("  rel L2", e, flush=True)
    assert e < 1e-5
