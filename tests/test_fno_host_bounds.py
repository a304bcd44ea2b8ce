"""CPU replay of the surrogate network's launch sequence: every kernel call of
a training step (4 blocks, the served configuration's depth) and of an
inference pass goes through the host-side operand checks of
atmosphericturbulencesimul_amd.fno.Ops with the library calls replaced by a
stand-in, so a wrong buffer or stride fails here instead of on the GPU.
Regression for a cache-key collision between block 1's pre-activation and the
projection's, which handed the GELU backward a buffer twice too large."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import atmosphericturbulencesimul_amd.fno as F  # noqa: E402


class _Lib:
    def __getattr__(self, name):
        return lambda *a: 0


class _Ctx:
    device = 0
    stream = None
    h = None
    L = _Lib()

    def _chk(self, rc):
        assert rc == 0

    def sync(self):
        pass


def _net(N=32, blocks=4, width=8, modes=4, hidden=16):
    return F.FNO(_Ctx(), N, 7, 2, width=width, modes=modes, blocks=blocks, hidden=hidden, device="cpu")


@pytest.mark.parametrize("blocks", [1, 2, 4])
def test_training_step_operands_in_bounds(blocks):
    net = _net(blocks=blocks)
    x = torch.zeros((7, 3, 32 * 32))
    t = torch.zeros((2, 3, 32 * 32))
    net.loss_and_grads(x, t, want_loss=False)
    net.adam_step()
    net.forward(x, keep=False)


def test_checks_catch_a_mismatched_buffer():
    ops = F.Ops(_Ctx())
    a = torch.zeros(64)
    with pytest.raises(ValueError):
        ops.gelu_bwd(a, torch.zeros(32), torch.zeros(32))
    with pytest.raises(ValueError):
        ops.gemm(8, 8, 8, 1, torch.zeros(63), 0, 0, 8, 1, torch.zeros(64), 0, 0, 8, 1, torch.zeros(64), 0, 0, 8, 1)
    ops.gemm(8, 8, 8, 1, torch.zeros(64), 0, 0, 8, 1, torch.zeros(64), 0, 0, 8, 1, torch.zeros(64), 0, 0, 8, 1)
