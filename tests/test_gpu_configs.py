"""GPU parity at the sizes of the measured workloads.

* 256^2 plain case: 3 layers at 0/5/10 km, Cn2 0.5/0.3/0.2, 2 cm pixels,
  1024^2 masters, 0.5 um, bilinear, 100 frames, every frame against the oracle.
* 1024^2 case: 7 layers at 0/0.5/1/2/4/8/16 km, Cn2 0.4/0.15/0.1/0.1/0.1/0.1/0.05,
  r0 0.12 m, 1 cm pixels, 4096^2 masters synthesised and normalised on the
  GPU, 5 wavelengths 0.5-1.65 um (reference rule at 0.5 um, block rule
  elsewhere), winds 5-40 m/s in seeded random directions; frames against the
  oracle fed the GPU's masters.
* 2048^2 headline (10 layers, 8 wavelengths, 8192^2 masters): size-independent
  properties -- every field keeps its energy N^2 through the layers and gaps
  (phase screens and the Fresnel transfer function are unitary), and without
  propagation the unwrapped phase equals the FP64 sum of the sampled screens
  from the oracle's sampler at 512 random pixels.
* 512^2 / 1024^2 / 2048^2 over one set of 8192^2 masters, 10 layers,
  8 wavelengths and 16 directions on a 4 x 4 grid at -1, -1/3, 1/3, 1 arcmin:
  energy of all 128 fields at every size, FP64 positions, and at 512^2 the
  two corner directions against the oracle.
* 4096^2, 6 layers (5 gaps), 3 wavelengths, 8192^2 GPU masters: one pass per
  gap against two passes per gap (relative L2 2e-6) and the energy N^2 of
  every field.

Bars as in test_gpu_sim: relative L2 1e-5 for fields and amplitudes; phase
relative L2 1e-5 on agreeing 2 pi decisions, at most 0.1 % disagreeing."""
import math

import numpy as np
import pytest

from tests.conftest import rel_l2
from tests.simcase import build_case, gpu_sim, oracle_sims
from tests.test_gpu_sim import _compare

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_plain_256_three_layers_100_frames(ctx, oracle):
    print("case: 256^2, M=1024, 3 layers, bilinear, 0.5 um, 100 frames", flush=True)
    c = build_case(oracle, N=256, Nout=256, M=1024, alts=[0.0, 5000.0, 10000.0], cn2=[0.5, 0.3, 0.2],
                   speeds=[10.0, 20.0, 30.0], dirs=np.radians([0.0, 90.0, 200.0]), lambdas=[0.5e-6],
                   lam_ref=0.5e-6, dx=0.02, dt=0.001, r0_m=0.15, L0=20.0, l0=0.01, bicubic=False)
    _compare(c, oracle_sims(oracle, c), gpu_sim(ctx, c), nframes=100)


def _gpu_masters_case(ctx, N, M, alts, cn2, speeds, dirs, lambdas, lam_ref, dx, r0, L0, l0, seed, fresnel=True,
                      bin_m=1000.0):
    from atmosphericturbulencesimul_amd.api import Simulation, superlayers
    L = len(alts)
    cn2 = np.asarray(cn2, np.float64) / np.sum(cn2)
    sa, sli = superlayers(alts, cn2, bin_m)
    x0 = np.full((1, L), 0.05 * M)
    y0 = np.full((1, L), 0.05 * M)
    vx = np.asarray(speeds) * np.cos(dirs) / dx
    vy = np.asarray(speeds) * np.sin(dirs) / dx
    lambdas = np.asarray(lambdas, np.float64)
    scoeff = np.array([[lam_ref / lam] * L for lam in lambdas])
    unwrap = [0 if abs(lam - lam_ref) < 1e-12 else 1 for lam in lambdas]
    sim = Simulation(ctx, N, N, M, sa, sli, lambdas, scoeff, unwrap, x0, y0, vx, vy, dx, 0.001, 0.0, True,
                     fresnel=fresnel)
    sim.synth_masters(seed, L0 / dx, l0 / dx)
    sim.normalize_masters(r0 / dx, cn2)
    return sim, dict(sa=sa, sli=sli, x0=x0, y0=y0, vx=vx, vy=vy, scoeff=scoeff, unwrap=unwrap, lambdas=lambdas)


def test_4096_one_pass_per_gap():
    # at 4096^2 the passes run the 16-value column kernel (4 columns per
    # workgroup) and the row kernel with 2 wavelengths per workgroup: one pass
    # per gap against two passes per gap (same Philox masters), and the
    # energy N^2 of every field
    import torch
    from atmosphericturbulencesimul_amd.api import Context
    print("case: 4096^2, M=8192, 6 layers (5 gaps), 3 wavelengths, 2 frames", flush=True)
    N, M, dx = 4096, 8192, 0.01
    alts = [0.0, 1500.0, 3000.0, 5000.0, 8000.0, 12000.0]
    cn2 = [0.35, 0.2, 0.15, 0.12, 0.1, 0.08]
    speeds = np.linspace(5.0, 30.0, 6)
    dirs = np.random.default_rng(0xA75 + 5000).uniform(0.0, 2 * math.pi, 6)
    lams = [0.5e-6, 1.1e-6, 2.2e-6]
    sims, ctxs = [], []
    for gp in (1, 2):
        cx = Context(0)
        cx.set_tuning("gap_passes", gp)
        ctxs.append(cx)
        sim, _ = _gpu_masters_case(cx, N, M, alts, cn2, speeds, dirs, lams, 0.5e-6, dx, 0.12, 20.0, 0.01,
                                   seed=0xA75 + 5000)
        sims.append(sim)
    assert sims[0].gaps() == 5
    fields = [torch.empty((1, 3, N, N), dtype=torch.complex64, device="cuda") for _ in sims]
    for k in range(2):
        outs = [s.frame(f) for s, f in zip(sims, fields)]
        e = (outs[0][1].double() ** 2).sum(dim=(-1, -2)).cpu().numpy().ravel() / (N * N)
        print(f"frame {k}: energy / N^2 - 1 =", np.abs(e - 1).max(), flush=True)
        assert np.abs(e - 1).max() < 1e-5
        for w in range(3):
            a, b = fields[0][0, w], fields[1][0, w]
            rel = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
            ra = float(torch.linalg.vector_norm(outs[0][1][0, w] - outs[1][1][0, w]) /
                       torch.linalg.vector_norm(outs[1][1][0, w]))
            print(f"frame {k} w {w}: field {rel:.2e} amp {ra:.2e}", flush=True)
            assert rel < 2e-6 and ra < 2e-6
    for s in sims:
        s.close()
    for cx in ctxs:
        cx.close()


def test_8192_one_gap_against_a_torch_fp64_step():
    # at 8192^2 (the row pass with its three table-twiddle stages and the
    # 2-column column pass): a one-gap frame -- the top layer's screen and the
    # x factor in the first row pass, the y factor in the column pass, the
    # ground layer's screen in the extraction -- against the same step in
    # torch FP64 (numpy's FFT conventions) from the frame's own sampled
    # screens, at the FP32 field bar 1e-5 (measured 7e-7; the screens come
    # back rounded to float); a wrong twiddle, transfer factor or layer order
    # is an O(1) error.
    from atmosphericturbulencesimul_amd.api import Context
    print("case: 8192^2, M=8192, 2 layers (1 gap), 0.5 um, 1 frame", flush=True)
    N, M, dx, lam, z = 8192, 8192, 0.01, 0.5e-6, 5000.0
    cx = Context(0)
    sim, _ = _gpu_masters_case(cx, N, M, [0.0, z], [0.6, 0.4], [10.0, 20.0], [0.3, 1.9], [lam], lam, dx, 0.12,
                               20.0, 0.01, seed=0xA75 + 5400)
    assert sim.gaps() == 1
    field = torch.empty((1, 1, N, N), dtype=torch.complex64, device="cuda")
    sim.frame(field)
    scr = torch.empty((1, 2, N * N), dtype=torch.float32, device="cuda")
    cx._chk(cx.L.atsim_sim_layer_screens(sim.h, scr.data_ptr()))
    cx.sync()
    v = scr.double().view(2, N, N)  # layer 0 at the ground, layer 1 at z
    kt = torch.fft.fftfreq(N, d=1.0 / N, device="cuda", dtype=torch.float64)  # signed frequency indices
    coeff = math.pi * z * lam / (dx * N) ** 2
    H = torch.exp(-1j * coeff * (kt[:, None] ** 2 + kt[None, :] ** 2))
    E = torch.fft.ifft2(torch.fft.fft2(torch.exp(1j * v[1])) * H) * torch.exp(1j * v[0])
    del H
    rel = float(torch.linalg.vector_norm(field[0, 0].to(torch.complex128) - E) / torch.linalg.vector_norm(E))
    print(f"field rel L2 vs torch FP64 {rel:.2e}", flush=True)
    assert rel < 1e-5
    sim.close()
    cx.close()


def test_1024_seven_layers_five_wavelengths(ctx, oracle):
    print("case: 1024^2, M=4096, 7 layers, 5 wavelengths, 2 frames", flush=True)
    N, M, dx = 1024, 4096, 0.01
    alts = [0.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0, 16000.0]
    cn2 = [0.4, 0.15, 0.1, 0.1, 0.1, 0.1, 0.05]
    speeds = np.linspace(5.0, 40.0, 7)
    dirs = np.random.default_rng(0xA75 + 2000).uniform(0.0, 2 * math.pi, 7)
    lams = [0.5e-6, 0.7e-6, 0.9e-6, 1.25e-6, 1.65e-6]
    sim, p = _gpu_masters_case(ctx, N, M, alts, cn2, speeds, dirs, lams, 0.5e-6, dx, 0.12, 20.0, 0.01,
                               seed=0xA75 + 2000)
    masters = np.stack([sim.master(l).cpu().numpy() for l in range(7)]).astype(np.float64)
    osim = oracle.Sim(N, N, masters, p["sa"], p["sli"], p["lambdas"], p["scoeff"], p["unwrap"], p["x0"][0],
                      p["y0"][0], p["vx"], p["vy"], dx, 0.001, 0.0, True)
    del masters

    class C_:
        pass
    c = C_()
    c.D, c.W, c.N, c.Nout = 1, 5, N, N
    _compare(c, [osim], sim, nframes=2)
    sim.close()


def test_headline_energy_and_sampled_phase(ctx, oracle):
    print("case: 2048^2, M=8192, 10 layers, 8 wavelengths, energy + 512-pixel phase", flush=True)
    N, M, dx, L = 2048, 8192, 0.01, 10
    alts = np.linspace(0.0, 20000.0, L)
    cn2 = np.exp(-alts / 2000.0)
    speeds = np.linspace(5.0, 40.0, L)
    dirs = np.arange(L) * 2.399963229728653
    lams = np.linspace(0.5e-6, 2.2e-6, 8)
    sim, p = _gpu_masters_case(ctx, N, M, alts, cn2, speeds, dirs, lams, 0.5e-6, dx, 0.12, 20.0, 0.01,
                               seed=0xA75 + 4000)
    for k in range(3):
        pha, amp = sim.frame()
        e = (amp.double() ** 2).sum(dim=(-1, -2)).cpu().numpy().ravel() / (N * N)
        print(f"frame {k}: energy / N^2 - 1 =", np.abs(e - 1).max(), flush=True)
        assert np.abs(e - 1).max() < 1e-5
    # without propagation: pha == sum_l scoeff_wl * screen_l at the frame's positions
    masters = [sim.master(l) for l in range(L)]
    sim2, _ = _gpu_masters_case(ctx, N, M, alts, cn2, speeds, dirs, lams, 0.5e-6, dx, 0.12, 20.0, 0.01,
                                seed=0xA75 + 4000, fresnel=False)
    for l in range(L):
        sim2.set_master(l, masters[l])
    for _ in range(2):
        pha2, amp2 = sim2.frame()
    assert (amp2 - 1.0).abs().max().item() < 1e-6
    xs, ys = sim2.positions()
    rng = np.random.default_rng(7)
    ii = rng.integers(0, N, 512)
    jj = rng.integers(0, N, 512)
    phi = np.zeros((len(lams), 512))
    for l in range(L):
        m = masters[l].cpu().numpy().astype(np.float64)
        v = np.array([oracle.sample(m, xs[0, l], ys[0, l], int(a), int(b), True) for a, b in zip(ii, jj)])
        phi += p["scoeff"][:, l][:, None] * v[None, :]
        del m
    g = pha2.cpu().numpy()[0][:, jj, ii]
    err = np.abs(g - phi).max()
    print("max |pha - Phi| =", err, "of", np.abs(phi).max(), flush=True)
    assert rel_l2(g, phi) < 1e-6
    sim.close()
    sim2.close()


def test_headline_against_the_oracle(ctx, oracle):
    # the headline workload itself: 2048^2, 10 layers (9 Fresnel gaps), 8
    # wavelengths 0.5-2.2 um with the reference rule at 0.5 um and the block
    # rule elsewhere, bicubic sampling, 8192^2 masters synthesised and
    # normalised on the GPU and handed to the FP64 oracle; two frames, every
    # field against the oracle at the bar of test_gpu_sim
    print("case: 2048^2, M=8192, 10 layers, 8 wavelengths, 2 frames against the oracle", flush=True)
    N, M, dx, L = 2048, 8192, 0.01, 10
    alts = np.linspace(0.0, 20000.0, L)
    cn2 = np.exp(-alts / 2000.0)
    speeds = np.linspace(5.0, 40.0, L)
    dirs = np.arange(L) * 2.399963229728653
    lams = np.linspace(0.5e-6, 2.2e-6, 8)
    sim, p = _gpu_masters_case(ctx, N, M, alts, cn2, speeds, dirs, lams, 0.5e-6, dx, 0.12, 20.0, 0.01,
                               seed=0xA75 + 4000)
    assert sim.gaps() == 9
    masters = np.stack([sim.master(l).cpu().numpy() for l in range(L)]).astype(np.float64)
    osim = oracle.Sim(N, N, masters, p["sa"], p["sli"], p["lambdas"], p["scoeff"], p["unwrap"], p["x0"][0],
                      p["y0"][0], p["vx"], p["vy"], dx, 0.001, 0.0, True)
    del masters

    class C_:
        pass
    c = C_()
    c.D, c.W, c.N, c.Nout = 1, 8, N, N
    worst = _compare(c, [osim], sim, nframes=2)
    print("headline worst against the FP64 oracle:", worst, flush=True)
    sim.close()


def test_mixed_sizes_sixteen_directions(ctx, oracle):
    # 512^2, 1024^2 and 2048^2 simulations (one per grid size) over the same
    # 10 masters of 8192^2, 8 wavelengths and 16 directions on a 4 x 4 grid
    # at -1, -1/3, 1/3, 1 arcmin (offset angle x altitude / dx per layer):
    # energy of all 128 fields at every size, FP64 positions of every
    # direction, and at 512^2 two corner directions against the oracle
    from atmosphericturbulencesimul_amd.api import Simulation, hip_copy, superlayers
    print("case: N in 512/1024/2048, M=8192, 10 layers, 8 wavelengths, 16 directions", flush=True)
    L, M, dx, dt = 10, 8192, 0.01, 0.001
    alts = 20000.0 * np.arange(L) / 9
    cn2 = np.exp(-alts / 2000.0)
    cn2 /= cn2.sum()
    speeds = np.linspace(5.0, 40.0, L)
    dirs = np.random.default_rng(0xA75 + 3000).uniform(0.0, 2 * math.pi, L)
    lams = np.linspace(0.5e-6, 2.2e-6, 8)
    ang = np.array([-1.0, -1.0 / 3, 1.0 / 3, 1.0]) * math.pi / (180 * 60)
    dgrid = [(gx, gy) for gy in ang for gx in ang]
    D = len(dgrid)
    x0 = np.array([[0.05 * M + gx * alts[l] / dx for l in range(L)] for gx, _ in dgrid])
    y0 = np.array([[0.05 * M + gy * alts[l] / dx for l in range(L)] for _, gy in dgrid])
    vx = speeds * np.cos(dirs) / dx
    vy = speeds * np.sin(dirs) / dx
    sa, sli = superlayers(alts, cn2, 1000.0)
    scoeff = np.array([[0.5e-6 / lam] * L for lam in lams])
    unwrap = [0] + [1] * 7
    sims = {}
    for N in (512, 1024, 2048):
        s = Simulation(ctx, N, N, M, sa, sli, lams, scoeff, unwrap, x0, y0, vx, vy, dx, dt, 0.0, True)
        if not sims:
            s.synth_masters(0xA75 + 3000, 20.0 / dx, 0.01 / dx)
            s.normalize_masters(0.12 / dx, cn2)
        else:
            first = next(iter(sims.values()))
            for l in range(L):
                hip_copy(ctx, s.master_ptr(l), first.master_ptr(l), 4 * M * M)
        sims[N] = s
    for N, s in sims.items():
        for k in range(2):
            pha, amp = s.frame()
            e = (amp.double() ** 2).sum(dim=(-1, -2)).cpu().numpy().ravel() / (N * N)
            print(f"N={N} frame {k}: energy / N^2 - 1 =", np.abs(e - 1).max(), flush=True)
            assert e.size == D * len(lams) and np.abs(e - 1).max() < 1e-5
        xg, yg = s.positions()
        # two frames of x += v dt from the offsets, FP64 on the host (no wrap this close to the start)
        xe, ye = x0.copy(), y0.copy()
        for _ in range(2):
            xe = xe + vx * dt
            ye = ye + vy * dt
        xe = np.where(xe < 0, xe + M, xe)
        ye = np.where(ye < 0, ye + M, ye)
        assert np.allclose(xg, xe, rtol=0, atol=1e-9) and np.allclose(yg, ye, rtol=0, atol=1e-9)
    # 512^2 against the oracle for the corner directions 0 and 15, all wavelengths
    s = sims[512]
    masters = np.stack([s.master(l).cpu().numpy() for l in range(L)]).astype(np.float64)
    osims = {d: oracle.Sim(512, 512, masters, sa, sli, lams, scoeff, unwrap, x0[d], y0[d], vx, vy, dx, dt, 0.0, True)
             for d in (0, D - 1)}
    field = torch.empty((D, len(lams), 512, 512), dtype=torch.complex64, device="cuda")
    s2 = Simulation(ctx, 512, 512, M, sa, sli, lams, scoeff, unwrap, x0, y0, vx, vy, dx, dt, 0.0, True)
    for l in range(L):
        hip_copy(ctx, s2.master_ptr(l), s.master_ptr(l), 4 * M * M)
    for k in range(2):
        pha_g, amp_g = s2.frame(field)
        fg, pg, ag = field.cpu().numpy(), pha_g.cpu().numpy(), amp_g.cpu().numpy()
        for d, os_ in osims.items():
            E_o, _, pha_o, amp_o = os_.frame()
            for w in range(len(lams)):
                assert rel_l2(fg[d, w], E_o[w]) < 1e-5
                assert rel_l2(ag[d, w], amp_o[w]) < 1e-5
                bad = np.abs(pg[d, w] - pha_o[w]) > 1.0
                assert bad.sum() <= 1e-3 * 512 * 512
                assert rel_l2(pg[d, w][~bad], pha_o[w][~bad]) < 1e-5
        print(f"oracle frame {k} ok", flush=True)
    del masters
    for x in list(sims.values()) + [s2]:
        x.close()
