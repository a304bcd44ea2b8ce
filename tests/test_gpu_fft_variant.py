"""The mixed-precision FFT variant: radix-16 stages as DFT-16 products on the
matrix cores with FP16 hi + lo operands and FP32 accumulation
(atsim_fft2d_variant / atsim_row_fft_variant, variant 1), against numpy's
FP64 FFT beside the FP32 Stockham transform (variant 0).

Bar: relative L2 <= 1e-5 for both variants (the FP32 bar of the path); the
measured errors of the two are printed side by side."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _field(shape, seed, smooth=False):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    if smooth:  # a unit-modulus field with a smooth random phase, like a propagated wavefront
        ph = np.cumsum(np.cumsum(rng.standard_normal(shape), -1), -2) * 0.01
        x = np.exp(1j * ph) * (1.0 + 0.1 * rng.standard_normal(shape))
    return x.astype(np.complex64)


@pytest.mark.parametrize("sign", [-1, 1])
def test_row_variant_vs_numpy(ctx, sign):
    n = 2048
    x = _field((n, n), 5 + sign)
    ref = np.fft.fft(x.astype(np.complex128), axis=-1) if sign < 0 else np.fft.ifft(x.astype(np.complex128), axis=-1) * n
    errs = []
    for variant in (0, 1):
        t = torch.tensor(x, device="cuda")
        torch.cuda.synchronize()
        ctx._chk(ctx.L.atsim_row_fft_variant(ctx.h, C.c_void_p(t.data_ptr()), n, n, sign, variant))
        ctx.sync()
        errs.append(rel_l2(t.cpu().numpy(), ref))
    print(f"row transforms, sign {sign}: FP32 Stockham {errs[0]:.2e}, FP16 hi+lo MFMA {errs[1]:.2e}", flush=True)
    assert max(errs) < 1e-5


@pytest.mark.parametrize("smooth", [False, True])
def test_2d_variant_vs_numpy(ctx, smooth):
    n = 2048
    x = _field((2, n, n), 7, smooth)
    errs = []
    for sign in (-1, 1):
        ref = np.fft.fft2(x.astype(np.complex128)) if sign < 0 else np.fft.ifft2(x.astype(np.complex128)) * n * n
        for variant in (0, 1):
            t = torch.tensor(x, device="cuda")
            torch.cuda.synchronize()
            ctx._chk(ctx.L.atsim_fft2d_variant(ctx.h, C.c_void_p(t.data_ptr()), n, 2, sign, variant))
            ctx.sync()
            errs.append(rel_l2(t.cpu().numpy(), ref))
    print(f"2D transforms (smooth={smooth}): forward {errs[0]:.2e} / {errs[1]:.2e}, inverse {errs[2]:.2e} / "
          f"{errs[3]:.2e} (FP32 Stockham / FP16 hi+lo MFMA)", flush=True)
    assert max(errs) < 1e-5
    # other sizes are refused
    from atmosphericturbulencesimul_amd._lib import AtsimError
    t = torch.zeros((1024, 1024), dtype=torch.complex64, device="cuda")
    with pytest.raises(AtsimError):
        ctx._chk(ctx.L.atsim_fft2d_variant(ctx.h, C.c_void_p(t.data_ptr()), 1024, 1, -1, 1))
