"""GPU parity of the frame loop (layer sampling, screen phase, Fresnel gaps,
extraction, both unwrap rules) against the FP64 oracle, through the C ABI.

Bar: complex field and amplitude within relative L2 1e-5; phase within
relative L2 1e-5 on pixels whose 2 pi decision agrees, with the number of
disagreeing decisions reported and bounded."""
import numpy as np
import pytest

from tests.conftest import rel_l2
from tests.simcase import build_case, gpu_sim, oracle_sims

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _compare(c, osims, gs, nframes, tol=1e-5, max_mismatch_frac=1e-3):
    field = torch.empty((c.D, c.W, c.N, c.N), dtype=torch.complex64, device="cuda")
    worst = {"E": 0.0, "amp": 0.0, "pha": 0.0, "mism": 0}
    for _ in range(nframes):
        pha_g, amp_g = gs.frame(field)
        fg = field.cpu().numpy()
        pg, ag = pha_g.cpu().numpy(), amp_g.cpu().numpy()
        for d, os_ in enumerate(osims):
            E_o, _, pha_o, amp_o = os_.frame()
            for w in range(c.W):
                worst["E"] = max(worst["E"], rel_l2(fg[d, w], E_o[w]))
                worst["amp"] = max(worst["amp"], rel_l2(ag[d, w], amp_o[w]))
                diff = pg[d, w] - pha_o[w]
                bad = np.abs(diff) > 1.0
                worst["mism"] = max(worst["mism"], int(bad.sum()))
                ok = ~bad
                worst["pha"] = max(worst["pha"], rel_l2(pg[d, w][ok], pha_o[w][ok]))
        xg, yg = gs.positions()
        for d, os_ in enumerate(osims):
            assert np.allclose(xg[d], os_.xpos, atol=0, rtol=1e-15)
            assert np.allclose(yg[d], os_.ypos, atol=0, rtol=1e-15)
    print("worst", worst)
    assert worst["E"] < tol
    assert worst["amp"] < tol
    assert worst["pha"] < tol
    assert worst["mism"] <= max_mismatch_frac * c.Nout * c.Nout
    return worst


def _cfg1_like(oracle, N=64, Nout=48, M=256, bicubic=True, **kw):
    # 3 layers at 0/5/10 km, Cn2 0.5/0.3/0.2, wind 10/20/30 m/s at 0/90/200 deg
    return build_case(oracle, N=N, Nout=Nout, M=M, alts=[0.0, 5000.0, 10000.0], cn2=[0.5, 0.3, 0.2],
                      speeds=[10.0, 20.0, 30.0], dirs=np.radians([0.0, 90.0, 200.0]),
                      lambdas=kw.pop("lambdas", [0.5e-6, 1.65e-6]), lam_ref=0.5e-6, dx=0.02, dt=0.001,
                      r0_m=0.15, L0=20.0, l0=0.01, bicubic=bicubic, **kw)


@pytest.mark.parametrize("bicubic", [True, False])
def test_frames_ref_and_sci_fields(ctx, oracle, bicubic):
    c = _cfg1_like(oracle, bicubic=bicubic, unwrap_mode=[0, 1])
    _compare(c, oracle_sims(oracle, c), gpu_sim(ctx, c), nframes=3)


def test_frames_two_directions_many_wavelengths(ctx, oracle):
    lams = np.linspace(0.5e-6, 2.2e-6, 9)
    c = _cfg1_like(oracle, N=64, Nout=64, M=128, lambdas=lams, unwrap_mode=[0] * 9,
                   directions=((0.0, 0.0), (2.9e-4, -2.9e-4)))
    _compare(c, oracle_sims(oracle, c), gpu_sim(ctx, c), nframes=2)


def test_frames_merged_superlayers(ctx, oracle):
    # layers closer than FRESNEL_PROPAGATION_BIN share a super-layer
    c = build_case(oracle, N=128, Nout=100, M=256, alts=[0.0, 500.0, 1000.0, 4000.0, 4300.0],
                   cn2=[0.4, 0.15, 0.1, 0.2, 0.15], speeds=[5, 10, 15, 20, 25],
                   dirs=[0.1, 1.0, 2.0, 3.0, 4.0], lambdas=[0.5e-6, 0.9e-6], lam_ref=0.5e-6, dx=0.01,
                   dt=0.001, r0_m=0.12, L0=20.0, l0=0.01, unwrap_mode=[0, 1], bin_m=1000.0)
    assert len(c.slayer_alt) < c.L
    gs = gpu_sim(ctx, c)
    assert gs.gaps() == len(c.slayer_alt) - 1
    _compare(c, oracle_sims(oracle, c), gs, nframes=2)


def test_phase_only_and_no_fresnel(ctx, oracle):
    c = _cfg1_like(oracle, unwrap_mode=[0, 1])
    _compare(c, oracle_sims(oracle, c, fresnel=False), gpu_sim(ctx, c, fresnel=False), nframes=2)
    c2 = _cfg1_like(oracle, unwrap_mode=[0, 0])
    os2 = oracle_sims(oracle, c2, amplitude=False)
    gs2 = gpu_sim(ctx, c2, amplitude=False)
    for _ in range(2):
        pg, ag = gs2.frame()
        _, _, po, ao = os2[0].frame(want_field=False)
        assert rel_l2(pg.cpu().numpy()[0], po) < 1e-6
        assert np.all(ag.cpu().numpy() == 1.0)


def test_wrap_around_master(ctx, oracle):
    # fast wind: positions wrap around the periodic master within a few frames
    c = build_case(oracle, N=64, Nout=64, M=64, alts=[0.0, 3000.0], cn2=[0.6, 0.4], speeds=[400.0, 900.0],
                   dirs=[2.5, -1.0], lambdas=[0.6e-6], lam_ref=0.6e-6, dx=0.05, dt=0.01, r0_m=0.2,
                   L0=10.0, l0=0.05)
    _compare(c, oracle_sims(oracle, c), gpu_sim(ctx, c), nframes=4)


def test_frame_graph_replays_match_eager(ctx, oracle):
    """graph mode (tuning key "graph"): the captured frame replayed over several
    frames, with other GPU work and synchronisations between the replays and a
    re-capture when the output pointers change, gives the eager frames"""
    from atmosphericturbulencesimul_amd.api import Context
    c = _cfg1_like(oracle, unwrap_mode=[0, 1])
    ctx_g = Context(0)
    ctx_g.set_tuning("graph", 1)
    assert ctx_g.get_tuning("graph") == 1
    ge = gpu_sim(ctx, c)
    gg = gpu_sim(ctx_g, c)
    field_e = torch.empty((c.D, c.W, c.N, c.N), dtype=torch.complex64, device="cuda")
    field_g = torch.empty_like(field_e)
    for k in range(6):
        use_field = k >= 4  # the last frames also write the field: new pointers, new capture
        pe, ae = ge.frame(field_e if use_field else None)
        pg, ag = gg.frame(field_g if use_field else None)
        x = torch.randn(1 << 20, device="cuda")
        _ = float(x.sum())  # unrelated work between replays
        torch.cuda.synchronize()
        assert torch.equal(ae, ag), k
        d = (pe - pg).abs()
        assert int((d > 1.0).sum()) <= 2, k  # block-sum atomics may order differently
        assert float(d[d <= 1.0].max()) < 1e-6, k
        if use_field:
            assert torch.equal(field_e, field_g), k
    xe, ye = ge.positions()
    xg, yg = gg.positions()
    assert np.array_equal(xe, xg) and np.array_equal(ye, yg)
    gg.close()
    ctx_g.close()


@pytest.mark.parametrize("separable", [True, False])
def test_layer_dependent_phase_multipliers(ctx, oracle, separable):
    # the reference recomputes Scoeff per layer (AtmosphericTurbulence.c:2661-2663):
    # multipliers that are, and are not, a wavelength factor times a layer factor
    sc = [[1.0, 1.0, 1.0], [0.303, 0.303, 0.303]] if separable else [[1.0, 1.0, 1.0], [0.301, 0.305, 0.312]]
    c = _cfg1_like(oracle, unwrap_mode=[0, 1], scoeff=sc)
    _compare(c, oracle_sims(oracle, c), gpu_sim(ctx, c), nframes=2)


@pytest.mark.parametrize("streams", [1, 2])
def test_extraction_from_geometric_phase_planes(oracle, streams):
    # factored multipliers (lambda_ref / lambda rows): the extraction reads
    # the two planes of the geometric phase made beside the passes (tuning
    # "extract_planes" 1) or samples every layer itself (2); amplitudes are
    # the same bits, phases agree to FP64 rounding, and both match the oracle
    from atmosphericturbulencesimul_amd.api import Context
    lams = [0.5e-6, 0.8e-6, 1.25e-6, 2.2e-6]
    c = _cfg1_like(oracle, N=128, Nout=100, M=256, lambdas=lams, unwrap_mode=[0, 1, 1, 1])
    print(f"case: 128^2 -> 100^2, 3 layers, 4 wavelengths, streams {streams}, 3 frames", flush=True)
    outs = []
    for mode in (1, 2):
        cx = Context(0)
        cx.set_tuning("extract_planes", mode)
        cx.set_tuning("streams", streams)
        assert cx.get_tuning("extract_planes") == mode
        gs = gpu_sim(cx, c)
        frames = []
        for _ in range(3):
            pha, amp = gs.frame()
            frames.append((pha.cpu().numpy().copy(), amp.cpu().numpy().copy()))
        gs.profile(True)
        gs.frame()
        assert gs.profile_read("geometry")[1] == (1 if mode == 1 else 0)
        gs.profile(False)
        outs.append(frames)
        gs.close()
        cx.close()
    for (p1, a1), (p2, a2) in zip(*outs):
        assert np.array_equal(a1, a2)
        d = np.abs(p1 - p2)
        assert int((d > 1.0).sum()) <= 2
        assert float(d[d <= 1.0].max()) < 1e-5
    cx = Context(0)
    cx.set_tuning("streams", streams)
    _compare(c, oracle_sims(oracle, c), gpu_sim(cx, c), nframes=2)
    cx.close()


@pytest.mark.parametrize("graph,streams", [(0, 2), (1, 1), (1, 2)])
def test_pipelined_frames(oracle, graph, streams):
    # tuning "pipeline" 1: frame k + 1's passes run on the simulation's own
    # stream beside frame k's extraction and fix-up (two buffer sets
    # alternate).  Every frame's outputs are the same bits as without it,
    # through alternating output buffers, a field output, a master replaced
    # between frames, a multi-simulation frame in the sequence and per-launch
    # profiling; both match the oracle.
    from atmosphericturbulencesimul_amd._lib import AtsimError
    from atmosphericturbulencesimul_amd.api import Context, sims_frame
    lams = [0.5e-6, 0.8e-6, 1.25e-6, 2.2e-6]
    c = _cfg1_like(oracle, N=128, Nout=100, M=256, lambdas=lams, unwrap_mode=[0, 1, 1, 1])
    print(f"case: 128^2 -> 100^2, 3 layers, 4 wavelengths, graph {graph}, streams {streams}, 9 frames", flush=True)
    runs = []
    for pipe in (0, 1):
        cx = Context(0)
        cx.set_tuning("graph", graph)
        cx.set_tuning("streams", streams)
        cx.set_tuning("pipeline", pipe)
        assert cx.get_tuning("pipeline") == pipe
        with pytest.raises(AtsimError):
            cx.set_tuning("pipeline", 2)
        assert cx.get_tuning("pipeline") == pipe
        gs = gpu_sim(cx, c)
        outs = [(torch.empty_like(gs.pha), torch.empty_like(gs.amp)) for _ in range(2)]
        field = torch.empty((c.D, c.W, c.N, c.N), dtype=torch.complex64, device="cuda")
        m1 = gs.master(1) * 0.5
        got = []
        for k in range(9):
            f = None
            if k == 3:
                gs.set_master(1, m1)
            if k == 5:
                sims_frame([gs], sync=False)
                pha, amp = gs.pha, gs.amp
            elif k == 6:
                gs.profile(True)
                pha, amp = gs.frame(sync=False, out=outs[k % 2])
                gs.profile(False)
            else:
                f = field if k == 4 else None
                pha, amp = gs.frame(f, sync=False, out=outs[k % 2])
            # copies ordered after the frame on the library's stream
            with torch.cuda.stream(cx.stream):
                got.append((pha.clone(), amp.clone(), None if f is None else f.clone()))
        cx.sync()
        torch.cuda.synchronize()
        runs.append([tuple(None if t is None else t.cpu().numpy() for t in g) for g in got])
        gs.close()
        cx.close()
    for k, (a, b) in enumerate(zip(*runs)):
        for x, y in zip(a, b):
            if x is not None:
                assert np.array_equal(x, y), f"frame {k}"
    cx = Context(0)
    cx.set_tuning("graph", graph)
    cx.set_tuning("streams", streams)
    cx.set_tuning("pipeline", 1)
    _compare(c, oracle_sims(oracle, c), gpu_sim(cx, c), nframes=3)
    cx.close()


def _block_sums_np(E, PHI, Nout):
    """sum cos / sin of PHI - arg E per (jj2, ii2) block over the centred
    Nout x Nout crop, blocks (long)(1.0 * i / Nout * xs) with xs = Nout / 20
    (AtmosphericTurbulence.c:3523-3625)"""
    N = E.shape[0]
    o = (N - Nout) // 2
    xs = Nout // 20
    d = PHI[o:o + Nout, o:o + Nout] - np.angle(E[o:o + Nout, o:o + Nout])
    k = (np.arange(Nout) / Nout * xs).astype(np.int64)
    bre = np.zeros((xs, xs))
    bim = np.zeros((xs, xs))
    np.add.at(bre, (k[:, None], k[None, :]), np.cos(d))
    np.add.at(bim, (k[:, None], k[None, :]), np.sin(d))
    return bre, bim


def test_block_sums_match_reference_sums(ctx, oracle):
    # two row segments per output row (Nout > 256), blocks of 20 and 21
    # pixels (Nout = 490, xs = 24), two directions, two science-rule fields
    c = _cfg1_like(oracle, N=512, Nout=490, M=1024, lambdas=[0.5e-6, 1.25e-6, 2.2e-6], unwrap_mode=[0, 1, 1],
                   directions=((0.0, 0.0), (2e-6, -1e-6)))
    osims = oracle_sims(oracle, c)
    gs = gpu_sim(ctx, c)
    field = torch.empty((c.D, c.W, c.N, c.N), dtype=torch.complex64, device="cuda")
    for _ in range(2):
        gs.frame(field)
        bs = gs.block_sums()
        for d, os_ in enumerate(osims):
            E_o, PHI_o, _, _ = os_.frame()
            for w in range(c.W):
                if c.unwrap[w] == 0:
                    assert not bs[d, w].any()
                    continue
                bre, bim = _block_sums_np(E_o[w], PHI_o[w], c.Nout)
                # a pixel left out or counted twice moves a sum by ~1
                err = max(np.abs(bs[d, w, :, :, 0] - bre).max(), np.abs(bs[d, w, :, :, 1] - bim).max())
                print("block sums", d, w, err, np.abs(bre).max())
                assert err < 2e-3


def test_context_closed_before_its_simulation(oracle):
    # the context closes the simulations still open on it; a later close or
    # garbage collection of the simulation must not touch the freed context
    from atmosphericturbulencesimul_amd.api import Context
    c = _cfg1_like(oracle)
    ctx2 = Context(0)
    gs = gpu_sim(ctx2, c)
    gs.frame()
    ctx2.close()
    assert gs.h is None
    gs.close()
    del gs


@pytest.mark.parametrize("N,cols", [(256, 8), (512, 8), (256, 16), (512, 16)])
def test_column_pass_32_elements_per_thread(oracle, N, cols):
    # column Fresnel pass with 32 elements per thread and float LDS exchanges,
    # 8 or 16 columns per workgroup
    from atmosphericturbulencesimul_amd.api import Context
    ctx2 = Context(0)
    ctx2.set_tuning("col_e", 32)
    ctx2.set_tuning("col_c", cols)
    assert ctx2.get_tuning("col_e") == 32
    c = _cfg1_like(oracle, N=N, Nout=N - 16, M=1024, unwrap_mode=[0, 1])
    _compare(c, oracle_sims(oracle, c), gpu_sim(ctx2, c), nframes=2)
    ctx2.close()


@pytest.mark.parametrize("Nout", [102, 121])
def test_block_rule_output_not_a_multiple_of_four(ctx, oracle, Nout):
    # block-rule unwrap over rows that the four-pixel path of the final pass
    # does not take (Nout % 4 != 0), blocks of 20 and 21 pixels
    c = _cfg1_like(oracle, N=128, Nout=Nout, M=256, unwrap_mode=[0, 1])
    _compare(c, oracle_sims(oracle, c), gpu_sim(ctx, c), nframes=2)


@pytest.mark.parametrize("r0_m", [0.08, 0.04])
def test_block_rule_in_strong_turbulence(ctx, oracle, r0_m):
    # strong turbulence at 0.5 and 0.8 um, both on the block rule: the oracle
    # finds about 160 (r0 8 cm) and 3 100 (4 cm) pixels per frame whose final
    # block count is not zero (the final phase is not arg E + 2 pi rint((PHI -
    # arg E) / 2 pi)); the GPU's phases match the oracle's on them too, and
    # the final update that visits only the blocks where a count can be
    # non-zero (tuning "block_apply" 1, the default) gives the bits of the
    # update over every pixel (2)
    from atmosphericturbulencesimul_amd.api import Context
    c = build_case(oracle, N=128, Nout=100, M=256, alts=[0.0, 4000.0, 12000.0], cn2=[0.5, 0.3, 0.2],
                   speeds=[10.0, 20.0, 30.0], dirs=np.radians([0.0, 90.0, 200.0]), lambdas=[0.5e-6, 0.8e-6],
                   lam_ref=0.5e-6, dx=0.02, dt=0.001, r0_m=r0_m, L0=20.0, l0=0.01, unwrap_mode=[1, 1])
    print(f"case: 128^2 -> 100^2, 3 layers, r0 {r0_m} m, block rule at 0.5 / 0.8 um, 3 frames", flush=True)
    cx = Context(0)
    cx.set_tuning("block_apply", 2)
    assert cx.get_tuning("block_apply") == 2 and ctx.get_tuning("block_apply") == 1
    gd = gpu_sim(cx, c)
    gs = gpu_sim(ctx, c)
    os_ = oracle_sims(oracle, c)[0]
    o = (c.N - c.Nout) // 2
    counted = 0
    for _ in range(3):
        pha_g = gs.frame()[0].cpu().numpy()
        assert np.array_equal(pha_g, gd.frame()[0].cpu().numpy())
        E, PHI, pha_o, _ = os_.frame()
        for w in range(c.W):
            Ec, Pc = E[w][o:o + c.Nout, o:o + c.Nout], PHI[w][o:o + c.Nout, o:o + c.Nout]
            a = np.arctan2(Ec.imag, Ec.real)
            counted += int((np.abs(pha_o[w] - (a + 2 * np.pi * np.rint((Pc - a) / (2 * np.pi)))) > 1.0).sum())
            bad = np.abs(pha_g[0, w] - pha_o[w]) > 1.0
            print(f"w {w}: {int(bad.sum())} disagreeing 2 pi decisions", flush=True)
            assert bad.sum() <= 5e-3 * c.Nout * c.Nout
            assert rel_l2(pha_g[0, w][~bad], pha_o[w][~bad]) < 1e-5
    print("pixels with a non-zero count in the oracle:", counted, flush=True)
    assert counted >= 300
    gs.close()
    gd.close()
    cx.close()


@pytest.mark.parametrize("streams,col_e", [(2, 16), (2, 32), (3, 16), (4, 16)])
def test_wavelength_sets_on_streams(oracle, streams, col_e):
    # wavelength sets on several streams give the one-stream frames exactly
    # (eager and graph-replayed frames) and match the oracle
    from atmosphericturbulencesimul_amd.api import Context
    lams = [0.5e-6, 0.8e-6, 1.25e-6, 1.65e-6, 2.2e-6]
    c = _cfg1_like(oracle, N=128, Nout=100, M=256, lambdas=lams, unwrap_mode=[0, 1, 1, 0, 1])
    outs = []
    for ns in (1, streams):
        cx = Context(0)
        cx.set_tuning("streams", ns)
        cx.set_tuning("col_e", col_e)
        cx.set_tuning("graph", 1)
        gs = gpu_sim(cx, c)
        frames = []
        for _ in range(4):
            pha, amp = gs.frame()
            frames.append((pha.cpu().numpy().copy(), amp.cpu().numpy().copy()))
        outs.append(frames)
        cx.close()
    for (p1, a1), (p2, a2) in zip(*outs):
        assert np.array_equal(p1, p2) and np.array_equal(a1, a2)
    cx = Context(0)
    cx.set_tuning("streams", streams)
    cx.set_tuning("col_e", col_e)
    _compare(c, oracle_sims(oracle, c), gpu_sim(cx, c), nframes=2)
    cx.close()


def _many_gaps_case(oracle, nl, N=128, Nout=100, M=256, bicubic=True, bin_m=0.0, directions=None):
    alts = [0.0, 700.0, 1500.0, 3000.0, 4500.0, 6500.0, 9000.0, 12000.0][:nl]
    cn2 = np.linspace(0.3, 0.05, nl)
    kw = {"directions": directions} if directions is not None else {}
    return build_case(oracle, N=N, Nout=Nout, M=M, alts=alts, cn2=list(cn2 / cn2.sum()),
                      speeds=list(np.linspace(5.0, 35.0, nl)), dirs=list(np.linspace(0.2, 5.5, nl)),
                      lambdas=[0.5e-6, 0.9e-6, 1.65e-6], lam_ref=0.5e-6, dx=0.01, dt=0.001, r0_m=0.12,
                      L0=20.0, l0=0.01, bicubic=bicubic, unwrap_mode=[0, 1, 1], bin_m=bin_m, **kw)


@pytest.mark.parametrize("nl,bicubic,bin_m,two_dirs,streams,N,col2_e", [
    (5, True, 0.0, False, 1, 128, 0),     # 4 gaps: the last pass is a row pass
    (6, True, 0.0, True, 2, 256, 0),      # 5 gaps: the last pass is a column pass; two directions, two stream sets
    (8, False, 0.0, False, 2, 256, 16),   # 7 gaps, bilinear, 16 elements per thread in the column passes
    (8, True, 2000.0, False, 1, 256, 0),  # merged super-layers: groups of several layers in row and column passes
    (7, True, 0.0, False, 1, 512, 0),
])
def test_one_pass_per_gap(oracle, nl, bicubic, bin_m, two_dirs, streams, N, col2_e):
    # one pass per gap (x / y order alternating) against a row and a column
    # pass per gap, and both against the oracle
    from atmosphericturbulencesimul_amd.api import Context
    dirs = ((0.0, 0.0), (2.9e-4, -1.5e-4)) if two_dirs else None
    c = _many_gaps_case(oracle, nl, N=N, Nout=N - 28, M=2 * N, bicubic=bicubic, bin_m=bin_m, directions=dirs)
    outs = []
    for gp in (1, 2):
        cx = Context(0)
        cx.set_tuning("gap_passes", gp)
        cx.set_tuning("streams", streams)
        cx.set_tuning("col2_e", col2_e)
        gs = gpu_sim(cx, c)
        if gp == 1:
            assert gs.gaps() >= 2
        field = torch.empty((c.D, c.W, c.N, c.N), dtype=torch.complex64, device="cuda")
        frames = []
        for _ in range(2):
            pha, amp = gs.frame(field)
            frames.append((field.cpu().numpy().copy(), pha.cpu().numpy().copy(), amp.cpu().numpy().copy()))
        outs.append(frames)
        cx.close()
    for (f1, p1, a1), (f2, p2, a2) in zip(*outs):
        for d in range(c.D):
            for w in range(c.W):
                assert rel_l2(f1[d, w], f2[d, w]) < 2e-6
                assert rel_l2(a1[d, w], a2[d, w]) < 2e-6
                ok = np.abs(p1[d, w] - p2[d, w]) < 1.0
                assert ok.mean() > 0.999
                assert rel_l2(p1[d, w][ok], p2[d, w][ok]) < 2e-6
    cx = Context(0)
    cx.set_tuning("streams", streams)
    _compare(c, oracle_sims(oracle, c), gpu_sim(cx, c), nframes=2)
    cx.close()


def test_tuning_change_between_graph_frames(oracle):
    # with graph replay, frames after a tuning change run the new kernel
    # sequence: the same frames as a context tuned that way from the start
    from atmosphericturbulencesimul_amd.api import Context
    c = _many_gaps_case(oracle, 6, N=256, Nout=228, M=512)
    runs = []
    for switch in (True, False):
        cx = Context(0)
        cx.set_tuning("graph", 1)
        if not switch:
            cx.set_tuning("gap_passes", 2)
            cx.set_tuning("streams", 1)
        gs = gpu_sim(cx, c)
        frames = []
        for k in range(6):
            if switch and k == 3:
                cx.set_tuning("gap_passes", 2)
                cx.set_tuning("streams", 1)
            pha, amp = gs.frame()
            frames.append((pha.cpu().numpy().copy(), amp.cpu().numpy().copy()))
        runs.append(frames)
        cx.close()
    for k in range(3, 6):
        assert np.array_equal(runs[0][k][0], runs[1][k][0]) and np.array_equal(runs[0][k][1], runs[1][k][1])
    # before the switch the frames came from one pass per gap: close, not equal
    assert rel_l2(runs[0][2][1], runs[1][2][1]) < 2e-6
    assert not np.array_equal(runs[0][2][1], runs[1][2][1])


@pytest.mark.parametrize("case", ["identity", "isolated_block", "branch_cut"])
def test_block_fixup_against_the_oracle(ctx, oracle, case):
    # the GPU fix-up heuristic on block sums from planted fields: nothing to
    # fix, an isolated block flipped in 2 rounds, and a straight branch cut
    # that runs into the 1000-round cap (AtmosphericTurbulence.c:3627-4262)
    n = 200
    y, x = np.mgrid[0:n, 0:n]
    if case == "identity":
        PHI = 25.0 * np.sin(x / 9.0) + 0.2 * y
        off = np.zeros((n, n))
    elif case == "isolated_block":
        PHI = 12.0 * np.cos(y / 30.0) + 0.05 * x
        off = np.full((n, n), 0.9 * np.pi)
        off[100:120, 100:120] = 1.1 * np.pi
    else:
        PHI = 12.0 * np.cos(y / 30.0)
        off = 0.6 * np.pi + 0.8 * np.pi * x / n
    E = np.exp(1j * (PHI + off))
    _, _, peak_o, it = oracle.extract_block(E, PHI, n)
    bre, bim = _block_sums_np(E, PHI, n)
    peak_g = ctx.block_fixup(np.stack([bre, bim], -1))
    print(case, "oracle rounds", it, "max |peak diff|", np.abs(peak_g - peak_o).max())
    assert np.allclose(peak_g, peak_o, rtol=0, atol=1e-9)


@pytest.mark.gpu
def test_block_counts_float_order_against_the_oracle(ctx, oracle):
    # first-round counts as the reference's float image accumulates them
    # (WFprecision = 0): every +-0.2 / +-0.15 rounded to float in the
    # reference's loop order; bit-exact against the oracle's float mode on
    # unwrapped peak grids where blocks get steps of both signs
    rng = np.random.default_rng(11)
    differs = 0
    for xs in (3, 7, 12, 25):
        for _ in range(10):
            peaks = rng.choice([-6.5, -3.3, 0.0, 3.3, 6.5], size=(xs, xs)).astype(np.float32).astype(np.float64)
            g = ctx.block_counts(peaks)
            _, ch32, _ = oracle.block_fix(peaks, prec_f32=True)
            _, ch64, _ = oracle.block_fix(peaks)
            assert np.array_equal(g.astype(np.float64), ch32), xs
            differs += int(np.sum(ch32 != ch64.astype(np.float32).astype(np.float64)))
    print("blocks whose float count differs from the rounded double count:", differs)
    assert differs > 0
