#!/usr/bin/env python3
"""Wavefront-frame throughput of the MI355X path on one GPU.

Workload (headline): 2048^2 propagation grid and output, 1 cm pixels,
8192^2 von Karman masters (L0 = 20 m, l0 = 1 cm) synthesised on the GPU from
seeded Philox noise, 10 layers at h_k = 20k/9 km with Cn2 ~ exp(-h / 2 km),
r0 = 0.12 m at 0.5 um, winds 5..40 m/s, 8 wavelengths 0.5..2.2 um, one
direction, Catmull-Rom sampling, Fresnel propagation between the 10
super-layers (9 gaps).  A step is one time step: every layer advanced, all 8
wavelength fields propagated and extracted (OPD + amplitude) = 8
wavefront-frames.

Printed JSON line: value = wavefront-frames/s with everything resident in
HBM; roofline of the dominant kernel from HIP events recorded around each of
its launches inside the timed region; cpu_baseline = the CPU oracle on a
bounded sample of the same workload (one wavelength per host thread), which
also gives the FP64 error of every full-size field; pcie_inclusive = the rate
with each frame handed to pinned host memory; per_wavelength_plain = the
same frames from one simulation per wavelength; pipelined = both rates again
with frame k + 1's passes beside frame k's extraction (tuning "pipeline"),
timed last.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "wavefront-frames/sec (λ×dir) at 2048², 10 layers; FFT HBM GB/s; err vs FP64"
HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 FP32 / FP64 instruction per
# 4 cycles on a 16-lane SIMD, 2.4 GHz
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 4


def headline_workload():
    L = 10
    alts = np.array([20000.0 * k / 9 for k in range(L)])
    cn2 = np.exp(-alts / 2000.0)
    cn2 /= cn2.sum()
    return dict(
        N=2048, Nout=2048, M=8192, L=L, W=8, D=1, dx=0.01, dt=0.001, r0=0.12, lam_ref=0.5e-6, L0=20.0, l0=0.01,
        alts=alts, cn2=cn2, speeds=np.linspace(5.0, 40.0, L),
        dirs=np.array([(2.399963229728653 * k) % (2 * math.pi) for k in range(L)]),
        lambdas=np.linspace(0.5e-6, 2.2e-6, 8), seed=0xA75 + 4000, bin_m=1000.0)


def make_sim(ctx, wl, M=None, lambdas=None, unwrap=None):
    from atmosphericturbulencesimul_amd.api import Simulation, superlayers
    M = M or wl["M"]
    lambdas = wl["lambdas"] if lambdas is None else np.asarray(lambdas)
    W = lambdas.size
    L = wl["L"]
    sa, sli = superlayers(wl["alts"], wl["cn2"], wl["bin_m"])
    x0 = np.full((1, L), 0.05 * M)
    y0 = np.full((1, L), 0.05 * M)
    vx = wl["speeds"] * np.cos(wl["dirs"]) / wl["dx"]
    vy = wl["speeds"] * np.sin(wl["dirs"]) / wl["dx"]
    # phase scale lambda_ref / lambda; the refractivity ratio N(h, lambda)/N(h, lambda_ref)
    # of the reference's atmosphere model is taken as 1 here
    scoeff = np.array([[wl["lam_ref"] / lam] * L for lam in lambdas])
    if unwrap is None:
        unwrap = [0 if abs(lam - wl["lam_ref"]) < 1e-12 else 1 for lam in lambdas]
    sim = Simulation(ctx, wl["N"], wl["Nout"], M, sa, sli, lambdas, scoeff, unwrap, x0, y0, vx, vy,
                     wl["dx"], wl["dt"], 0.0, True)
    return sim, dict(sa=sa, sli=sli, x0=x0, y0=y0, vx=vx, vy=vy, scoeff=scoeff, unwrap=np.asarray(unwrap))


def read_pass_table(fname, column):
    """{pass: value} from the two-column table headed `| pass | <column> |` of a
    committed rocprof PMC summary under profiles/"""
    p = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(p):
        return None
    out, inside = {}, False
    for line in open(p):
        cells = [c.strip() for c in line.strip().strip("|").split("|")] if line.startswith("|") else None
        if cells is None:
            inside = False
            continue
        if cells == ["pass", column]:
            inside = True
            continue
        if inside and len(cells) == 2 and not cells[1].startswith("-"):
            try:
                out[cells[0]] = float(cells[1])
            except ValueError:
                pass
    return out or None


def read_traffic():
    """HBM bytes per launch of each pass from the committed FETCH_SIZE / WRITE_SIZE summary"""
    return read_pass_table("hbm_traffic.md", "bytes_per_launch")


def read_valu():
    """VALU wave instructions per launch of each pass from the committed SQ_INSTS_VALU summary"""
    return read_pass_table("valu_headline_2048_10L_8lam.md", "wave_instructions_per_launch")


def _fftw_backend():
    """Name of the FFT library a reference build on this host would use: FFTW
    (libfftw3f.so.3, src/fft/fft.c:52) if the host has it, else none"""
    import ctypes
    for name in ("libfftw3f.so.3", "libfftw3f.so"):
        try:
            ctypes.CDLL(name)
            return name
        except OSError:
            pass
    return None


def _oracle_case(wl, masters_gpu, M):
    """host float32 masters cut to M^2 from the GPU's, and the simulation set-up"""
    ms = np.stack([m[:M, :M].cpu().numpy() for m in masters_gpu]).astype(np.float32)
    return ms


def fp64_error(ctx, wl, ms, frames=1, threads=0):
    """FP64 error of every full-size field: the GPU and the FP64 oracle run the
    same frames of the 2048^2, 10-layer, bicubic, Fresnel loop for all 8
    wavelengths (reference rule at 0.5 um, block rule elsewhere) over masters
    cut to 2048^2 from the GPU's; one oracle simulation per wavelength, each on
    its own host thread"""
    import torch
    from concurrent.futures import ThreadPoolExecutor
    from oracle import pyoracle as po
    M = ms.shape[-1]
    L, W = wl["L"], wl["W"]
    sim, p = make_sim(ctx, wl, M=M)
    for l in range(L):
        sim.set_master(l, ms[l])
    ms64 = ms.astype(np.float64)
    osims = [po.Sim(wl["N"], wl["Nout"], ms64, p["sa"], p["sli"], [wl["lambdas"][w]], p["scoeff"][w:w + 1],
                    [int(p["unwrap"][w])], p["x0"][0], p["y0"][0], p["vx"], p["vy"], wl["dx"], wl["dt"], 0.0, True)
              for w in range(W)]
    field = torch.empty((1, W, wl["N"], wl["N"]), dtype=torch.complex64, device="cuda")
    gpu = []
    for _ in range(frames):
        pha_g, amp_g = sim.frame(field)
        gpu.append((field.cpu().numpy()[0], pha_g.cpu().numpy()[0], amp_g.cpu().numpy()[0]))
    sim.close()
    threads = threads or min(W, 16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        outs = list(ex.map(lambda o: [o.frame() for _ in range(frames)], osims))
    del ms64
    errs = dict(field=0.0, amp=0.0, opd=0.0, unwrap_mismatch=0, scint_index=0.0)
    rl = lambda a, b: float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))

    def si(a):
        I = a.astype(np.float64) ** 2
        return float(I.var() / I.mean() ** 2)

    for w in range(W):
        for k in range(frames):
            E_o, _, pha_o, amp_o = outs[w][k]
            fg, pg, ag = gpu[k][0][w], gpu[k][1][w], gpu[k][2][w]
            errs["field"] = max(errs["field"], rl(fg, E_o[0]))
            errs["amp"] = max(errs["amp"], rl(ag, amp_o[0]))
            bad = np.abs(pg - pha_o[0]) > 1.0
            errs["unwrap_mismatch"] = max(errs["unwrap_mismatch"], int(bad.sum()))
            errs["opd"] = max(errs["opd"], rl(pg[~bad], pha_o[0][~bad]))
            s_o = si(amp_o[0])
            errs["scint_index"] = max(errs["scint_index"], abs(si(ag) - s_o) / s_o)
    errs["case"] = (f"all {W} fields (0.5-2.2 um), 2048^2, 10 layers, {frames} frame(s), vs FP64 oracle (max "
                    "over fields and frames; OPD over pixels whose 2 pi decision agrees; scintillation index "
                    "sigma_I^2 relative error)")
    return errs


def cpu_baseline(wl, ms, frames=1, threads=0):
    """The reference's own structure on the host cores: one run per science
    wavelength, each carrying the reference-wavelength field beside its own
    (AtmosphericTurbulence.c:2672, 2685), float storage with FP64 per-pixel
    arithmetic (WFprecision = 0, AT.c:1472-1615), bicubic sampling, block rule
    for the science field.  The 8 distinct wavefront-frames of a step (7
    science wavelengths + the reference one) take the 7 two-field runs.
    Timed on one core (the reference's path is single-threaded) and with the
    runs spread over the host cores."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import pyoracle as po
    from atmosphericturbulencesimul_amd.api import superlayers
    L, W = wl["L"], wl["W"]
    M = ms.shape[-1]
    sa, sli = superlayers(wl["alts"], wl["cn2"], wl["bin_m"])
    lam_ref = wl["lam_ref"]
    sci = [lam for lam in wl["lambdas"] if abs(lam - lam_ref) > 1e-12]
    x0 = np.full(L, 0.05 * M)
    vx = wl["speeds"] * np.cos(wl["dirs"]) / wl["dx"]
    vy = wl["speeds"] * np.sin(wl["dirs"]) / wl["dx"]
    ms64 = ms.astype(np.float64)

    def run(lam):
        return po.Sim(wl["N"], wl["Nout"], ms64, sa, sli, [lam_ref, lam],
                      np.array([[1.0] * L, [lam_ref / lam] * L]), [0, 1], x0, x0, vx, vy, wl["dx"], wl["dt"],
                      0.0, True, prec_f32=True)

    # one core: one two-field run
    r1 = run(sci[-1])
    t0 = time.perf_counter()
    for _ in range(frames):
        r1.frame(want_field=False)
    t_one = (time.perf_counter() - t0) / frames  # seconds per two-field run frame
    del r1
    # all cores: the 7 runs side by side
    threads = threads or min(len(sci), 16, os.cpu_count() or 1)
    runs = [run(lam) for lam in sci]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda r: [r.frame(want_field=False) for _ in range(frames)], runs))
    t_all = time.perf_counter() - t0
    del runs, ms64
    fftw = _fftw_backend()
    cpu_model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                cpu_model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": round(frames * W / t_all, 4),
        "unit": "wavefront-frames/s",
        "cores": threads,
        "kind": "port",
        "single_core": round(W / (len(sci) * t_one), 4),
        "fft_backend": "oracle radix-2 FFT (FP64 arithmetic, fields stored as float)" +
                       ("" if fftw is None else f"; {fftw} present but not used"),
        "fftw_present": fftw is not None,
        "host": {"nproc": os.cpu_count(), "cpu_model": cpu_model},
        "sample": f"{frames} frame(s) of the 2048^2, 10-layer, 9-gap bicubic workload as the reference runs it: "
                  f"{len(sci)} runs (one per science wavelength), each propagating the 0.5 um reference field "
                  f"beside its own, float storage with FP64 per-pixel arithmetic (WFprecision = 0), masters "
                  f"{M}^2 cut from the GPU's; {W} distinct wavefront-frames per step; all cores: the runs on "
                  f"{threads} threads, {t_all:.1f} s; one core: one run, {t_one:.1f} s per frame",
    }


def mixed_sizes(ctx, wl, owner, steps):
    """The batched uneven-size workload: 512^2, 1024^2 and 2048^2 simulations
    (16 directions on a 4 x 4 grid at -1, -1/3, 1/3, 1 arcmin, 8 wavelengths,
    10 layers) over the headline's 8192^2 masters.  Rates in wavefront-frames/s
    (384 per step): all three sizes in one call (atsim_sims_frame, each size
    on its own stream, eager launches), the three simulations one after the
    other, and the plain per-wavelength launch (one field per simulation;
    timed on one direction per size and scaled to the 16)."""
    import torch
    from atmosphericturbulencesimul_amd.api import Simulation, superlayers, sims_frame
    sizes = (512, 1024, 2048)
    L, M = wl["L"], wl["M"]
    sa, sli = superlayers(wl["alts"], wl["cn2"], wl["bin_m"])
    ang = np.array([-1.0, -1.0 / 3, 1.0 / 3, 1.0]) * math.pi / (180 * 60)
    dirs = [(tx, ty) for ty in ang for tx in ang]
    D, W = len(dirs), wl["W"]
    vx = wl["speeds"] * np.cos(wl["dirs"]) / wl["dx"]
    vy = wl["speeds"] * np.sin(wl["dirs"]) / wl["dx"]
    scoeff = np.array([[wl["lam_ref"] / lam] * L for lam in wl["lambdas"]])
    unwrap = [0 if abs(lam - wl["lam_ref"]) < 1e-12 else 1 for lam in wl["lambdas"]]

    def make(N, dset, lams, sc, uw):
        x0 = np.array([[0.05 * M + tx * a / wl["dx"] for a in wl["alts"]] for tx, _ in dset])
        y0 = np.array([[0.05 * M + ty * a / wl["dx"] for a in wl["alts"]] for _, ty in dset])
        sim = Simulation(ctx, N, N, M, sa, sli, lams, sc, uw, x0, y0, vx, vy, wl["dx"], wl["dt"], 0.0, True)
        sim.use_masters(owner)
        return sim

    sims = [make(N, dirs, wl["lambdas"], scoeff, unwrap) for N in sizes]
    F = D * W * len(sizes)
    for _ in range(2):
        sims_frame(sims, sync=False)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        sims_frame(sims, sync=False)
    ctx.sync()
    t_multi = (time.perf_counter() - t0) / steps
    for s in sims:
        s.frame(sync=False)
        s.frame(sync=False)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        for s in sims:
            s.frame(sync=False)
    ctx.sync()
    t_seq = (time.perf_counter() - t0) / steps
    for s in sims:
        s.close()
    del sims
    torch.cuda.empty_cache()
    # plain per-wavelength launches: one direction, one field per simulation
    t_plain = 0.0
    for N in sizes:
        one = [make(N, dirs[:1], wl["lambdas"][w:w + 1], scoeff[w:w + 1], [unwrap[w]]) for w in range(W)]
        for s in one:
            s.frame(sync=False)
            s.frame(sync=False)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            for s in one:
                s.frame(sync=False)
        ctx.sync()
        t_plain += D * (time.perf_counter() - t0) / steps
        for s in one:
            s.close()
    return {
        "workload": "512^2 + 1024^2 + 2048^2, 16 directions (4 x 4 grid, +-1/3 and +-1 arcmin), 8 wavelengths, "
                    "10 layers, 8192^2 masters shared",
        "wavefront_frames_per_step": F,
        "one_call": round(F / t_multi, 2),
        "one_after_another": round(F / t_seq, 2),
        "per_wavelength_plain": round(F / t_plain, 2),
        "one_call_vs_per_wavelength": round(t_plain / t_multi, 2),
        "unit": "wavefront-frames/s",
        "steps": steps,
        "how": "one_call: atsim_sims_frame (each size on its own stream, eager launches); one_after_another: "
               "the three simulations' frames in sequence; per_wavelength_plain: one field per simulation, "
               "timed on one direction per size and scaled to 16",
    }


def fft_variants(ctx, sim, reps=10):
    """The 2048^2 2D FFT of the headline's fields (a frame's 8 final fields) in
    the two arithmetic variants of atsim_fft2d_variant: 0 = FP32 Stockham,
    1 = radix-16 stages as FP16 hi + lo DFT-16 products on the matrix cores.
    Relative L2 error of each against numpy's FP64 FFT of the same fields,
    and device time of the row transforms of all 8 x 2048 rows and of the
    whole 2D transform (events on the library stream)."""
    import ctypes as C
    import torch
    n = sim.N
    field = torch.empty((sim.D, sim.W, n, n), dtype=torch.complex64, device="cuda")
    sim.frame(field)
    x = field.reshape(-1, n, n).contiguous()
    nb = x.shape[0]
    ref = np.fft.fft2(x.cpu().numpy().astype(np.complex128))
    st = ctx.stream
    out = {}
    for variant in (0, 1):
        y = x.clone()
        torch.cuda.synchronize()
        ctx._chk(ctx.L.atsim_fft2d_variant(ctx.h, C.c_void_p(y.data_ptr()), n, nb, -1, variant))
        ctx.sync()
        err = float(np.linalg.norm((y.cpu().numpy() - ref).ravel()) / np.linalg.norm(ref.ravel()))
        y = x.clone()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            ctx._chk(ctx.L.atsim_row_fft_variant(ctx.h, C.c_void_p(y.data_ptr()), n, n * nb, -1, variant))
        e1.record(st)
        e1.synchronize()
        t_rows = e0.elapsed_time(e1) / reps
        e0.record(st)
        for _ in range(reps):
            ctx._chk(ctx.L.atsim_fft2d_variant(ctx.h, C.c_void_p(y.data_ptr()), n, nb, -1, variant))
        e1.record(st)
        e1.synchronize()
        t_2d = e0.elapsed_time(e1) / reps
        name = "fp32_stockham" if variant == 0 else "fp16_hilo_mfma"
        out[name] = {"rel_l2_vs_fp64": err, "row_pass_ms": round(t_rows, 4),
                     "row_pass_gbs": round(16.0 * n * n * nb / (t_rows * 1e-3) / 1e9, 1),
                     "fft2d_ms": round(t_2d, 4)}
    out["case"] = (f"forward 2D FFT of the {nb} final 2048^2 fields of a headline frame; the FP16 variant's "
                   "2D transform runs rows, transpose, rows, transpose")
    return out


def shm_stream(frames, workdir):
    """The 2048^2 headline streamed as the reference runs it: `atsim mkwfs 0.7 0`
    (the science field at 0.7 um beside the 0.5 um reference field, 10
    layers, 8192^2 masters) writing every science frame's OPD [m] and
    amplitude to the shared-memory images <prefix>pha / <prefix>amp (with the
    TIME keyword and the semaphores posted) and the per-time-span cubes in
    host memory, 100 frames per time span.  Three runs in one directory: 100
    frames (synthesises and caches the masters), 100 frames (cached masters),
    `frames` frames; the streaming rate is (frames - 100) over the time the
    third run takes beyond the second."""
    import subprocess
    cli = os.path.join(ROOT, "atmosphericturbulencesimul_amd", "atsim")
    os.makedirs(workdir, exist_ok=True)
    L = 10
    alts = [20000.0 * k / 9 for k in range(L)]
    prof = "# alt cn2 speed dir L0 l0 sigmaW Lwind\n" + "".join(
        f"{a:.1f} {math.exp(-a / 2000.0):.6e} {5.0 + 35.0 * k / (L - 1):.3f} "
        f"{(2.399963229728653 * k) % (2 * math.pi):.6f} 20.0 0.01 0 0\n" for k, a in enumerate(alts))
    with open(os.path.join(workdir, "prof.txt"), "w") as f:
        f.write(prof)

    def conf(nspan):
        return ("TURBULENCE_REF_WAVEL 0.5\nTURBULENCE_SEEING 0.859436\nTURBULENCE_PROF_FILE prof.txt\n"
                "ZENITH_ANGLE 0.0\nWFOUTPUT 2\nSHM_OUTPUT 0\nMAKE_SWAVEFRONT 1\nSWF_WRITE2DISK 0\n"
                "SHM_SOUTPUT 1\nSHM_SPREFIX atsimstream_\nSHM_SOUTPUTM 1\nPUPIL_SCALE 0.01\nWFsize 2048\n"
                f"WFTIME_STEP 0.001\nTIME_SPAN 0.1\nNB_TSPAN {nspan}\nWF_RAW_SIZE 2048\nMASTER_SIZE 8192\n"
                "WAVEFRONT_AMPLITUDE 1\nFRESNEL_PROPAGATION 1\nFRESNEL_PROPAGATION_BIN 1000.0\n"
                "ATSIM_NOISE_SEED 19233\n")

    env = dict(os.environ, ATSIM_SHM_DIR=workdir)
    times = []
    for nspan in (1, 1, frames // 100):
        with open(os.path.join(workdir, "WFsim.conf"), "w") as f:
            f.write(conf(nspan))
        t0 = time.perf_counter()
        p = subprocess.run([cli, "mkwfs", "0.7", "0"], cwd=workdir, env=env, capture_output=True, text=True,
                           timeout=900)
        times.append(time.perf_counter() - t0)
        print(f"mkwfs {nspan * 100} frames: {times[-1]:.2f} s, rc {p.returncode}", file=sys.stderr, flush=True)
        if p.returncode != 0:
            raise RuntimeError("mkwfs failed: " + p.stdout[-2000:] + p.stderr[-2000:])
    nfr = (frames // 100) * 100
    rate = (nfr - 100) / (times[2] - times[1])
    shm = os.path.join(workdir, "atsimstream_pha.im.shm")
    return {"metric": "shared-memory streamed frames/s (2048^2, 10 layers, science 0.7 um + 0.5 um reference field)",
            "value": round(rate, 2), "unit": "frames/s", "frames": nfr, "wall_s": [round(t, 2) for t in times],
            "shm_image_bytes": os.path.getsize(shm) if os.path.exists(shm) else None,
            "how": "atsim mkwfs 0.7 0 (compat frame loop: two fields per step, double-buffered device-to-host "
                   "copies, shared-memory OPD and amplitude per frame, 100-frame host cubes); rate = extra "
                   "frames over extra time between a 100-frame and the full run on cached masters"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / FP64 error sample")
    ap.add_argument("--cpu-frames", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0, help="host threads of the CPU baseline (0 = one per "
                    "wavelength, at most 16)")
    ap.add_argument("--row-mw", type=int, default=0, help="wavelengths per row-pass workgroup (0 = default)")
    ap.add_argument("--col-c", type=int, default=0, help="columns per column-pass workgroup (0 = default)")
    ap.add_argument("--streams", type=int, default=0, help="Fresnel passes of this many wavelength sets on as many "
                    "streams (0 = default, 2; 1 = one stream)")
    ap.add_argument("--col-e", type=int, default=0, help="elements per thread of the column pass (0 = default, "
                    "16, 32)")
    ap.add_argument("--col2-e", type=int, default=0, help="elements per thread of the column passes that apply "
                    "a screen (0 = default, 16, 32)")
    ap.add_argument("--gap-passes", type=int, default=0, help="passes over the fields per Fresnel gap (0 = "
                    "default, 1: x / y order alternating; 2: a row and a column pass)")
    ap.add_argument("--autotune", type=int, default=1, help="1: launch shapes picked by the library's autotuner "
                    "(atsim_sim_autotune) before the timed frames; 0: built-in defaults or the flags above")
    ap.add_argument("--graph", type=int, default=1, help="1: frames replay a captured HIP graph of their kernels "
                    "(0: eager launches)")
    ap.add_argument("--pipeline", type=int, default=0, help="1: frame k + 1's Fresnel passes run beside frame k's "
                    "extraction and block fix-up in the timed frames (two buffer sets; 0: frames one after the other)")
    ap.add_argument("--pipelined", type=int, default=1, help="1 (with --pipeline 0): time the same frames and the "
                    "per-wavelength ones pipelined at the end of the run ('pipelined' in the line)")
    ap.add_argument("--host-copy-steps", type=int, default=20,
                    help="extra frames each followed by a synchronous copy of OPD + amplitude to pinned host "
                         "memory (the PCIe-inclusive rate of a host hand-over; 0 = skip)")
    ap.add_argument("--plain-steps", type=int, default=20,
                    help="frames of the per-wavelength comparison (one simulation per wavelength; 0 = skip)")
    ap.add_argument("--fft-variants", type=int, default=1, help="1: time and check the FP32 and FP16-MFMA 2D FFT "
                    "variants on the headline's fields")
    ap.add_argument("--mixed-steps", type=int, default=5,
                    help="steps of the uneven-size workload (512/1024/2048, 16 directions; 0 = skip)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one plain device-to-device copy of this size first (calibrates rocprof byte counters)")
    ap.add_argument("--shm-stream", type=int, default=0,
                    help="instead of the headline line: stream this many frames (a multiple of 100) of the "
                         "2048^2 workload through atsim mkwfs to shared memory and print that rate")
    ap.add_argument("--shm-dir", default="", help="working directory of --shm-stream (default: a temporary "
                    "directory, removed afterwards; it holds 2.7 GB of cached masters)")
    args = ap.parse_args()
    if args.shm_stream > 0:
        import tempfile
        if args.shm_dir:
            print(json.dumps(shm_stream(args.shm_stream, os.path.abspath(args.shm_dir))))
        else:
            with tempfile.TemporaryDirectory(prefix="atsim_stream_") as d:
                print(json.dumps(shm_stream(args.shm_stream, d)))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the wavefront path is built and measured on a single MI355X "
              "(independent (lambda, direction) shards would split across GPUs with no exchange); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from atmosphericturbulencesimul_amd.api import Context

    wl = headline_workload()
    ctx = Context(0)
    if args.row_mw:
        ctx.set_tuning("row_mw", args.row_mw)
    if args.col_c:
        ctx.set_tuning("col_c", args.col_c)
    if args.col_e:
        ctx.set_tuning("col_e", args.col_e)
    if args.streams:
        ctx.set_tuning("streams", args.streams)
    if args.gap_passes:
        ctx.set_tuning("gap_passes", args.gap_passes)
    if args.col2_e:
        ctx.set_tuning("col2_e", args.col2_e)
    if args.graph:
        ctx.set_tuning("graph", args.graph)
    if args.calib_copy_mib > 0:
        n = args.calib_copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    sim, setup = make_sim(ctx, wl)
    t0 = time.perf_counter()
    sim.synth_masters(wl["seed"], wl["L0"] / wl["dx"], wl["l0"] / wl["dx"])
    r0_meas = sim.normalize_masters(wl["r0"] / wl["dx"], wl["cn2"])
    t_setup = time.perf_counter() - t0
    ctx.set_tuning("pipeline", args.pipeline)
    tune = None
    if args.autotune:
        t0 = time.perf_counter()
        tune = sim.autotune(frames=8)
        tune["search_s"] = round(time.perf_counter() - t0, 3)
    G = sim.gaps()
    F = wl["W"] * wl["D"]

    for _ in range(args.warmup):
        sim.frame(sync=False)
    ctx.sync()
    # kernel durations from HIP events around every launch of the profiled
    # frames (recorded on the library's stream); with graph replay the timed
    # frames below run without events and a separate profiled run supplies
    # the per-kernel averages of the same kernels
    if not args.graph:
        sim.profile(True)
    torch.cuda.synchronize()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.frame(sync=False)
    ctx.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if args.graph:
        sim.profile(True)
        for _ in range(min(args.steps, 20)):
            sim.frame(sync=False)
        ctx.sync()
    prof = {k: sim.profile_read(k) for k in ("row", "col", "row2", "col2", "phase", "geometry", "extract", "block")}
    sim.profile(False)
    pcie = None
    if args.host_copy_steps > 0:
        # frames handed to pinned host memory: two output sets, the copy of
        # frame k (on the library's copy stream) overlapping frame k + 1
        L_, h_ = ctx.L, ctx.h
        nb = sim.pha.numel() * 4
        outs = [(torch.empty_like(sim.pha), torch.empty_like(sim.amp)) for _ in range(2)]
        host = [(torch.empty(tuple(sim.pha.shape), dtype=torch.float32, pin_memory=True),
                 torch.empty(tuple(sim.amp.shape), dtype=torch.float32, pin_memory=True)) for _ in range(2)]
        torch.cuda.synchronize()
        ctx.sync()

        def handover(k):
            b = k & 1
            if k >= 2:
                ctx._chk(L_.atsim_stream_wait_copy(h_, 2 * b))
                ctx._chk(L_.atsim_stream_wait_copy(h_, 2 * b + 1))
            sim.frame(sync=False, out=outs[b])
            ctx._chk(L_.atsim_copy_async(h_, 2 * b, host[b][0].data_ptr(), outs[b][0].data_ptr(), nb))
            ctx._chk(L_.atsim_copy_async(h_, 2 * b + 1, host[b][1].data_ptr(), outs[b][1].data_ptr(), nb))
            if k >= 1:  # the host side of frame k - 1 is complete here
                ctx._chk(L_.atsim_copy_wait(h_, 2 * (1 - b)))
                ctx._chk(L_.atsim_copy_wait(h_, 2 * (1 - b) + 1))

        for k in range(4):
            handover(k)
        for b in range(4):
            ctx._chk(L_.atsim_copy_wait(h_, b))
        t0 = time.perf_counter()
        for k in range(args.host_copy_steps):
            handover(k)
        for b in range(4):
            ctx._chk(L_.atsim_copy_wait(h_, b))
        el = time.perf_counter() - t0
        pcie = {"value": round(args.host_copy_steps * F / el, 2), "unit": "wavefront-frames/s",
                "steps": args.host_copy_steps, "bytes_to_host_per_step": 2 * nb,
                "how": "OPD and amplitude of every frame copied to pinned host memory on a second stream, "
                       "double-buffered so that frame k's copy overlaps frame k + 1"}
    per_lambda = None
    if args.plain_steps > 0:
        # the same wavefront-frames from one simulation per wavelength (one
        # field per launch), the reference's per-wavelength structure
        sims1 = []
        for w in range(wl["W"]):
            s1, _ = make_sim(ctx, wl, lambdas=wl["lambdas"][w:w + 1], unwrap=[int(setup["unwrap"][w])])
            for l in range(wl["L"]):
                s1.set_master(l, sim.master(l))
            sims1.append(s1)
        for s1 in sims1:
            s1.frame(sync=False)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            for s1 in sims1:
                s1.frame(sync=False)
        ctx.sync()
        el = time.perf_counter() - t0
        v1 = args.plain_steps * wl["W"] / el
        per_lambda = {"value": round(v1, 2), "unit": "wavefront-frames/s", "steps": args.plain_steps,
                      "how": "one simulation per wavelength, one field per launch, same masters and kernels"}

    variants = fft_variants(ctx, sim) if args.fft_variants else None
    mixed = None
    if args.mixed_steps > 0:
        mixed = mixed_sizes(ctx, wl, sim, args.mixed_steps)
    pipelined = None
    if args.pipelined and not args.pipeline:
        # the same frames with frame k + 1's passes beside frame k's
        # extraction and fix-up (tuning "pipeline"), last, so that its extra
        # stream cannot change how the measurements above share the GPU's
        # hardware queues
        ctx.set_tuning("pipeline", 1)

        def rate(sims, steps):
            for s_ in sims:
                for _ in range(3):
                    s_.frame(sync=False)
            ctx.sync()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                for s_ in sims:
                    s_.frame(sync=False)
            ctx.sync()
            torch.cuda.synchronize()
            return time.perf_counter() - t0

        el = rate([sim], args.steps)
        pipelined = {"value": round(args.steps * F / el, 2), "unit": "wavefront-frames/s", "steps": args.steps,
                     "ms_per_step": round(el / args.steps * 1e3, 4)}
        if per_lambda is not None:
            el = rate(sims1, args.plain_steps)
            pipelined["per_wavelength_plain"] = round(args.plain_steps * wl["W"] / el, 2)
        ctx.set_tuning("pipeline", 0)
    if per_lambda is not None:
        for s1 in sims1:
            s1.close()

    N = wl["N"]
    kinds = ("row", "col", "row2", "col2", "phase")
    # Algorithmic bytes of every Fresnel pass of a frame, by kind: one c64
    # read + write of every field (the first pass creates the field: write
    # only) and, per direction, one f32 read of the footprint of each layer
    # whose screen a row pass applies, or one f64 read of the sampled phase
    # plane of each layer a column pass applies (the phase kernel reads the
    # f32 footprint and writes the f64 plane).  Layer groups top -> bottom as
    # the library forms them; the last group is applied by the extraction.
    sli = list(np.asarray(setup["sli"]))
    L = wl["L"]
    lcount = []
    for l in range(L - 1, -1, -1):
        if l == L - 1 or sli[l + 1] != sli[l]:
            lcount.append(1)
        else:
            lcount[-1] += 1
    fused = ctx.get_tuning("gap_passes") == 1 and N <= 4096 and G >= 2
    field_rw, field_w, foot = 16.0 * N * N * F, 8.0 * N * N * F, 4.0 * N * N * wl["D"]
    passes = []  # (kind, bytes)
    if fused:
        nslots = sum(lcount[k] for k in range(1, G, 2))
        passes.append(("phase", 3.0 * foot * nslots))
        for k in range(G + 1):
            rows = k % 2 == 0
            if k == 0:
                passes.append(("row", field_w + foot * lcount[0]))
            elif k == G:
                passes.append(("row" if rows else "col", field_rw))
            elif rows:
                passes.append(("row2", field_rw + foot * lcount[k]))
            else:
                passes.append(("col2", field_rw + 2.0 * foot * lcount[k]))
    else:
        for g in range(G):
            passes.append(("row", (field_w if g == 0 else field_rw) + foot * lcount[g]))
            passes.append(("col", field_rw))
    kb = {k: sum(b for kk, b in passes if kk == k) for k in kinds}
    kn = {k: sum(1 for kk, _ in passes if kk == k) for k in kinds}
    bytes_per_launch = {k: kb[k] / kn[k] for k in kinds if kn[k]}
    avg = {k: ms / n for k, (ms, n) in prof.items() if n}
    dom = max(kinds, key=lambda k: prof[k][0])
    dom_bytes = bytes_per_launch[dom]
    achieved = dom_bytes / (avg[dom] * 1e-3) / 1e9
    fft_ms = sum(prof[k][0] for k in kinds)
    fft_gbs = sum(bytes_per_launch[k] * prof[k][1] for k in kinds if prof[k][1]) / (fft_ms * 1e-3) / 1e9
    value = args.steps * F / elapsed
    # whole-frame algorithmic bytes per wavefront-frame: the passes above,
    # the extraction's c64 read of the field, an f32 OPD and amplitude write;
    # with the factored geometric phase (lambda_ref / lambda multipliers, the
    # default path) also the sweep's f32 read of every layer's footprint, its
    # two f64 planes written and the extraction's read of them
    frame_bytes = (sum(b for _, b in passes) + 8.0 * N * N * F) / F + 8.0 * wl["Nout"] ** 2
    if ctx.get_tuning("extract_planes") == 1:
        frame_bytes += (foot * L + 32.0 * N * N * wl["D"]) / F
    c2e = ctx.get_tuning("col2_e") or (32 if 256 <= N <= 2048 else 16)
    ce = ctx.get_tuning("col_e") or 16
    cc = ctx.get_tuning("col_c") or (8 if N == 2048 else 16)
    kernel_names = {"row": f"k_row_fresnel<{N}> (x pass: screen phase, FFT, transfer function, IFFT)",
                    "col": (f"k_col_split<{N}, 8, 32>" if ce == 32 and 256 <= N <= 2048 and cc in (8, 16)
                            else f"k_col<{N}, 1, 0, {cc}>") + " (y pass: FFT, transfer function, IFFT)",
                    "row2": f"k_row_fresnel<{N}, false, true> (x pass: FFT, transfer function, IFFT, screen "
                            "phase, FFT, transfer function, IFFT)",
                    "col2": (f"k_col_screen_split<{N}, 8, 32>" if c2e == 32 else f"k_col_screen<{N}>") +
                            " (y pass: FFT, transfer function, IFFT, screen phase, FFT, transfer function, IFFT)"}
    # the committed counter summaries were taken with the kernels named above
    # (the default shapes); another pick of the autotuner has no summary
    default_shapes = c2e == 32 and not (ce == 32) and cc == 8
    traffic = None
    tr = read_traffic()
    if tr and dom in tr and default_shapes:
        traffic = tr[dom]
    valu = None
    vi = read_valu()
    if vi and dom in vi and default_shapes:
        rate = vi[dom] / (avg[dom] * 1e-3)
        valu = {"wave_instructions_per_launch": vi[dom], "achieved_per_s": round(rate, -6),
                "peak_per_s": VALU_PEAK_WAVE_INSTR_PER_S, "frac": round(rate / VALU_PEAK_WAVE_INSTR_PER_S, 4)}

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "wavefront-frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (Philox-seeded noise, masters synthesised on the GPU)",
        "config": {
            "workload": "2048^2 grid, 10 layers (9 Fresnel gaps), 8 wavelengths 0.5-2.2 um, 1 direction, "
                        "8192^2 masters, bicubic, block unwrap on 7 science wavelengths",
            "N": N, "Nout": wl["Nout"], "M": wl["M"], "layers": wl["L"], "gaps": int(G),
            "wavelengths": wl["W"], "directions": wl["D"], "wavefront_frames_per_step": F,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_names[dom],
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": dom_bytes,
            "launches_per_frame": kn[dom],
            # the fused passes are bound by VALU issue (FP32 FFT arithmetic)
            # before HBM: wave instructions per launch (SQ_INSTS_VALU) over
            # the live launch duration, against the issue peak
            "valu": valu,
            # the limiter the counters show: VALU issue when its fraction of
            # the issue peak exceeds the kernel's fraction of HBM peak
            "measured_limiter": (None if valu is None else
                                 ("valu_issue" if valu["frac"] > achieved / HBM_PEAK_GBS else "hbm")),
            "avg_launch_ms": round(avg[dom], 5),
        },
        "fft_hbm_gbs": round(fft_gbs, 1),
        "frame_roofline": {"bytes_per_wavefront_frame": frame_bytes,
                           "achieved_gbs": round(frame_bytes * value / 1e9, 1),
                           "frac": round(frame_bytes * value / 1e9 / HBM_PEAK_GBS, 4)},
        "kernel_avg_ms": {k: round(v, 5) for k, v in avg.items()},
        "pcie_inclusive": pcie,
        "per_wavelength_plain": per_lambda,
        "batched_vs_per_wavelength": round(value / per_lambda["value"], 2) if per_lambda else None,
        "mixed_sizes": mixed,
        "pipelined": pipelined,
        "fft_variants": variants,
        "setup_s": round(t_setup, 3),
        "tuning": {"row_mw": ctx.get_tuning("row_mw"), "col_c": ctx.get_tuning("col_c"),
                   "col_e": ctx.get_tuning("col_e"), "streams": ctx.get_tuning("streams"),
                   "gap_passes": ctx.get_tuning("gap_passes"), "col2_e": ctx.get_tuning("col2_e"),
                   "graph": ctx.get_tuning("graph"), "pipeline": ctx.get_tuning("pipeline")},
        "autotune": tune,
        "r0_masters_px_before_norm": [round(float(v), 3) for v in r0_meas],
    }
    if not args.no_cpu:
        masters = [sim.master(l) for l in range(wl["L"])]
        sim.close()
        ms = _oracle_case(wl, masters, wl["N"])
        del masters
        line["err_vs_fp64"] = fp64_error(ctx, wl, ms, frames=args.cpu_frames, threads=args.cpu_threads)
        line["cpu_baseline"] = cpu_baseline(wl, ms, frames=args.cpu_frames, threads=args.cpu_threads)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
