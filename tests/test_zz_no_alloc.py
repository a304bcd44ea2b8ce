"""The allocation-free Fresnel entry point against atsim_fresnel, through the C ABI.

atsim_fresnel_noalloc keeps its transfer-factor and twiddle tables in a fixed
device array of the library instead of allocating them; its results must be
those of atsim_fresnel (relative L2 <= 1e-5, the floating-point bar of the
Fresnel tests), on two calls in a row, the second on a smaller input.
"""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _field(torch, n, nb, seed):
    g = torch.Generator().manual_seed(seed)
    re = torch.randn((nb, n, n), generator=g)
    im = torch.randn((nb, n, n), generator=g)
    return torch.complex(re, im).to(torch.complex64).cuda().contiguous()


def test_noalloc_fresnel_matches_fresnel(ctx):
    import torch
    L = ctx.L
    vp = lambda t: C.c_void_p(t.data_ptr())
    for n, nb, dx, z, lam, seed in [(1024, 3, 0.01, 2222.2, 1.65e-6, 1), (256, 2, 0.02, 5000.0, 0.5e-6, 2)]:
        print(f"noalloc fresnel n={n} nbatch={nb} z={z} lambda={lam}", flush=True)
        src = _field(torch, n, nb, seed)
        ref = torch.empty_like(src)
        out = torch.empty_like(src)
        torch.cuda.synchronize()
        assert L.atsim_fresnel(ctx.h, vp(src), vp(ref), n, nb, dx, z, lam) == 0
        assert L.atsim_fresnel_noalloc(ctx.h, vp(src), vp(out), n, nb, dx, z, lam) == 0
        assert L.atsim_sync(ctx.h) == 0
        r, o = ref.cpu().numpy(), out.cpu().numpy()
        err = rel_l2(o, r)
        print(f"  rel L2 vs atsim_fresnel {err:.2e}, bit-equal {np.array_equal(o, r)}", flush=True)
        assert err <= 1e-5
    # in place, and sizes beyond the fixed array are refused
    inp = _field(torch, 512, 1, 3)
    ref = torch.empty_like(inp)
    torch.cuda.synchronize()
    assert L.atsim_fresnel(ctx.h, vp(inp), vp(ref), 512, 1, 0.01, 1000.0, 1e-6) == 0
    assert L.atsim_fresnel_noalloc(ctx.h, vp(inp), vp(inp), 512, 1, 0.01, 1000.0, 1e-6) == 0
    assert L.atsim_sync(ctx.h) == 0
    assert rel_l2(inp.cpu().numpy(), ref.cpu().numpy()) <= 1e-5
    assert L.atsim_fresnel_noalloc(ctx.h, vp(inp), vp(inp), 16384, 1, 0.01, 1000.0, 1e-6) == -4
    assert b"exceeds the maximum" in L.atsim_error(ctx.h)
