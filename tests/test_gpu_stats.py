"""The per-frame OPD / scintillation statistics stream (atsim_sim_stats) on the GPU.

Per field of a frame: OPD piston <pha> lambda / 2 pi and rms, mean intensity
<amp^2> and scintillation index <I^2> / <I>^2 - 1 over the Nout^2 output
pixels.  Checked against the same statistics of the GPU's own outputs in
numpy FP64 (relative 1e-9: only the summation order differs) and of the FP64
oracle's frames (relative 1e-5 for piston, rms and mean intensity, the bar of
the outputs they come from; 1e-3 for the scintillation index, a variance of
a quantity near 1, whose relative error is that of the intensities divided by
the index).
"""
import numpy as np
import pytest

from tests.simcase import build_case, gpu_sim, oracle_sims

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _np_stats(pha, amp, lam):
    I = amp.astype(np.float64) ** 2
    p = pha.astype(np.float64)
    mi = I.mean()
    return np.array([p.mean() * lam / (2 * np.pi), p.std() * lam / (2 * np.pi), mi, (I ** 2).mean() / mi ** 2 - 1])


def _case(oracle, **kw):
    return build_case(oracle, N=kw.pop("N", 256), Nout=kw.pop("Nout", 200), M=kw.pop("M", 512),
                      alts=[0.0, 4000.0, 10000.0], cn2=[0.5, 0.3, 0.2], speeds=[10.0, 20.0, 30.0],
                      dirs=np.radians([0.0, 90.0, 200.0]), lambdas=kw.pop("lambdas", [0.5e-6, 0.9e-6, 1.65e-6]),
                      lam_ref=0.5e-6, dx=0.01, dt=0.001, r0_m=0.1, L0=20.0, l0=0.01, **kw)


@pytest.mark.parametrize("graph", [0, 1])
def test_stats_against_outputs_and_oracle(ctx, oracle, graph):
    c = _case(oracle, unwrap_mode=[0, 1, 1], directions=((0.0, 0.0), (2.9e-4, -1.5e-4)))
    ctx.set_tuning("graph", graph)
    try:
        gs = gpu_sim(ctx, c)
        gs.stats_enable(True)
        osims = oracle_sims(oracle, c)
        worst = np.zeros(4)
        for frame in range(4):
            print(f"stats graph={graph} frame {frame}", flush=True)
            pha, amp = gs.frame()
            st = gs.stats()
            pg, ag = pha.cpu().numpy(), amp.cpu().numpy()
            for d, os_ in enumerate(osims):
                _, _, pha_o, amp_o = os_.frame()
                for w in range(c.W):
                    own = _np_stats(pg[d, w], ag[d, w], c.lambdas[w])
                    assert np.allclose(st[d, w], own, rtol=1e-9, atol=0), (d, w, st[d, w], own)
                    ref = _np_stats(pha_o[w], amp_o[w], c.lambdas[w])
                    worst = np.maximum(worst, np.abs(st[d, w] - ref) / np.abs(ref))
        print("worst relative error vs the FP64 oracle (piston, rms, <I>, sigma_I^2):", worst)
        assert np.all(worst[:3] < 1e-5) and worst[3] < 1e-3
        # the index is not degenerate: the 10 km layer scintillates at 0.5 um
        assert st[0, 0, 3] > 1e-3
    finally:
        ctx.set_tuning("graph", 0)


def test_stats_toggle_and_phase_only(ctx, oracle):
    c = _case(oracle, N=128, Nout=128, M=256, lambdas=[0.5e-6], unwrap_mode=[0])
    gs = gpu_sim(ctx, c, amplitude=False)
    with pytest.raises(Exception):
        gs.stats()  # not enabled yet
    gs.stats_enable(True)
    pha, amp = gs.frame()
    st = gs.stats()
    assert np.allclose(st[0, 0], _np_stats(pha.cpu().numpy()[0, 0], amp.cpu().numpy()[0, 0], c.lambdas[0]),
                       rtol=1e-9, atol=1e-300)
    assert st[0, 0, 2] == 1.0 and st[0, 0, 3] == 0.0  # phase only: amp = 1
    gs.stats_enable(False)
    gs.frame()
    assert np.array_equal(gs.stats(), st)  # frames without statistics leave the last ones
