"""Python front end of libatsim.so (GPU only, no fallback).

The functions mirror the reference's operators for the wavefront path:

    Context.fft2d            do2dfft / do2dffti      (src/fft/fft.c:1298-1315)
    Context.permut           permut                  (src/fft/fft.c:570-792)
    Context.fresnel          Fresnel_propagate_wavefront        (WFpropagate.c:68-152)
    Context.fresnel_tf_init  Init_Fresnel_propagate_wavefront   (WFpropagate.c:166-261)
    Context.fresnel_tf_apply Fresnel_propagate_wavefront1       (WFpropagate.c:271-330)
    Context.fresnel_cube     Fresnel_propagate_cube             (WFpropagate.c:337-401)
    Context.make_master      make_master_turbulence_screen      (AtmosphericTurbulence.c:643-855)
    Simulation.frame         one frame of make_AtmosphericTurbulence_wavefront_series
                             (AtmosphericTurbulence.c:2561-4262)

Tensors are torch CUDA tensors (complex64 fields, float32 screens/outputs).
"""
import ctypes as C
import weakref

import numpy as np

from . import _lib
from ._lib import AtsimError, SimDesc


def _torch():
    import torch
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Context:
    def __init__(self, device=0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise AtsimError("no GPU visible: the wavefront path runs on the GPU only")
        self.L = _lib.lib()
        self.device = device
        h = C.c_void_p()
        rc = self.L.atsim_ctx_create(device, C.byref(h))
        if rc != 0:
            raise AtsimError(f"atsim_ctx_create failed ({rc})")
        self.h = h
        self._torch_stream = torch.cuda.ExternalStream(self.L.atsim_stream(h), device=f"cuda:{device}")
        self._sims = weakref.WeakSet()  # simulations on this context, closed before it

    def close(self):
        if getattr(self, "h", None):
            for s in list(getattr(self, "_sims", ())):
                s.close()
            self.L.atsim_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- util
    def _chk(self, rc):
        if rc != 0:
            raise AtsimError(f"atsim error {rc}: {self.L.atsim_error(self.h).decode()}")

    def sync(self):
        self._chk(self.L.atsim_sync(self.h))

    @property
    def stream(self):
        """torch view of the library's HIP stream"""
        return self._torch_stream

    def _prep(self, t):
        """make pending torch work visible to the library stream"""
        torch = _torch()
        torch.cuda.current_stream(t.device).synchronize()
        return t

    def block_fixup(self, bsum):
        """(xs, xs) block peaks after the fix-up heuristic, from (xs, xs, 2) block sums"""
        b = np.ascontiguousarray(bsum, np.float64)
        xs = b.shape[0]
        assert b.shape == (xs, xs, 2)
        out = np.zeros((xs, xs))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self._chk(self.L.atsim_block_fixup(self.h, dp(b), xs, dp(out)))
        return out

    def block_counts(self, peaks):
        """(xs, xs) float32 first-round counts of the fix-up for (xs, xs) block peaks"""
        b = np.ascontiguousarray(peaks, np.float64)
        xs = b.shape[0]
        assert b.shape == (xs, xs)
        out = np.zeros((xs, xs), np.float32)
        self._chk(self.L.atsim_block_counts(self.h, b.ctypes.data_as(C.POINTER(C.c_double)), xs,
                                            out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_tuning(self, key, value):
        self._chk(self.L.atsim_set_tuning(self.h, key.encode(), int(value)))

    def get_tuning(self, key):
        return self.L.atsim_get_tuning(self.h, key.encode())

    def device_ok(self):
        return bool(self.L.atsim_device_ok(self.h))

    # ---------------------------------------------------------------- FFT
    def fft2d(self, t, sign=-1):
        """in place, unnormalised; t complex64 (..., n, n)"""
        assert t.is_contiguous() and t.shape[-1] == t.shape[-2]
        n = t.shape[-1]
        nb = t.numel() // (n * n)
        self._prep(t)
        self._chk(self.L.atsim_fft2d(self.h, _ptr(t), n, nb, sign))
        self.sync()
        return t

    def fft2d_nm(self, t, sign=-1):
        """in place, unnormalised, any (..., n1, n0) complex64 shape"""
        assert t.is_contiguous()
        n1, n0 = t.shape[-2], t.shape[-1]
        nb = t.numel() // (n0 * n1)
        self._prep(t)
        self._chk(self.L.atsim_fft2d_nm(self.h, _ptr(t), n0, n1, nb, sign))
        self.sync()
        return t

    def fresnel_nm(self, t_in, dx, z, lam, out=None):
        """Fresnel_propagate_wavefront on any (..., n1, n0) complex64 shape"""
        torch = _torch()
        assert t_in.is_contiguous() and t_in.dtype == torch.complex64
        n1, n0 = t_in.shape[-2], t_in.shape[-1]
        nb = t_in.numel() // (n0 * n1)
        out = torch.empty_like(t_in) if out is None else out
        self._prep(t_in)
        self._chk(self.L.atsim_fresnel_nm(self.h, _ptr(t_in), _ptr(out), n0, n1, nb, dx, z, lam))
        self.sync()
        return out

    def permut(self, t):
        assert t.is_contiguous()
        n1, n0 = t.shape[-2], t.shape[-1]
        nb = t.numel() // (n0 * n1)
        self._prep(t)
        self._chk(self.L.atsim_permut(self.h, _ptr(t), n0, n1, nb))
        self.sync()
        return t

    # ---------------------------------------------------------------- Fresnel
    def fresnel(self, t_in, dx, z, lam, out=None):
        torch = _torch()
        assert t_in.is_contiguous() and t_in.dtype == torch.complex64
        n = t_in.shape[-1]
        nb = t_in.numel() // (n * n)
        out = torch.empty_like(t_in) if out is None else out
        self._prep(t_in)
        self._chk(self.L.atsim_fresnel(self.h, _ptr(t_in), _ptr(out), n, nb, dx, z, lam))
        self.sync()
        return out

    def fresnel_tf_init(self, n, dx, z, lam, fpmaskrad):
        torch = _torch()
        Ct = torch.empty((n, n), dtype=torch.complex64, device=f"cuda:{self.device}")
        self._chk(self.L.atsim_fresnel_tf_init(self.h, _ptr(Ct), n, dx, z, lam, fpmaskrad))
        self.sync()
        return Ct

    def fresnel_tf_apply(self, t_in, Ct, out=None):
        torch = _torch()
        n = t_in.shape[-1]
        nb = t_in.numel() // (n * n)
        out = torch.empty_like(t_in) if out is None else out
        self._prep(t_in)
        self._chk(self.L.atsim_fresnel_tf_apply(self.h, _ptr(t_in), _ptr(out), _ptr(Ct), n, nb))
        self.sync()
        return out

    def fresnel_cube(self, t_in, dx, zstart, zend, nz, lam):
        torch = _torch()
        n = t_in.shape[-1]
        amp = torch.empty((nz, n, n), dtype=torch.float32, device=t_in.device)
        pha = torch.empty_like(amp)
        self._prep(t_in)
        self._chk(self.L.atsim_fresnel_cube(self.h, _ptr(t_in), _ptr(amp), _ptr(pha), n, dx, zstart, zend, nz, lam))
        self.sync()
        return amp, pha

    # ---------------------------------------------------------------- noise / screens
    def noise(self, n, seed, stream):
        torch = _torch()
        t = torch.empty(n, dtype=torch.complex64, device=f"cuda:{self.device}")
        self._chk(self.L.atsim_noise_c64(self.h, _ptr(t), n, seed, stream))
        self.sync()
        return t

    def make_master(self, noise, outerscale_px, innerscale_px, rlim=0.0, want_second=True):
        torch = _torch()
        size = noise.shape[-1]
        o1 = torch.empty((size, size), dtype=torch.float32, device=noise.device)
        o2 = torch.empty_like(o1) if want_second else None
        c1 = C.c_double()
        c2 = C.c_double()
        self._prep(noise)
        self._chk(self.L.atsim_make_master(self.h, _ptr(noise), size, outerscale_px, innerscale_px, rlim,
                                           _ptr(o1), _ptr(o2), C.byref(c1), C.byref(c2)))
        self.sync()
        return o1, o2, c1.value, c2.value

    def strfn(self, s, a, b):
        out = C.c_double()
        self._prep(s)
        self._chk(self.L.atsim_strfn(self.h, _ptr(s), s.shape[-1], a, b, C.byref(out)))
        return out.value

    def r0_measure(self, s, dpix=50):
        out = C.c_double()
        self._prep(s)
        self._chk(self.L.atsim_r0_measure(self.h, _ptr(s), s.shape[-1], dpix, C.byref(out)))
        return out.value

    def scale(self, s, factor):
        self._prep(s)
        self._chk(self.L.atsim_scale(self.h, _ptr(s), s.numel(), factor))
        self.sync()
        return s


def hip_copy(ctx, dst, src, nbytes):
    """device/host copy through the library's runtime (hipMemcpyDefault)"""
    ctx._chk(ctx.L.atsim_copy(ctx.h, C.c_void_p(dst), C.c_void_p(src), nbytes))


def superlayers(alt, cn2, bin_m):
    """atsim_superlayers: super-layer altitudes and layer -> super-layer index"""
    L_ = _lib.lib()
    alt = np.ascontiguousarray(alt, np.float64)
    cn2 = np.ascontiguousarray(cn2, np.float64)
    n = alt.size
    sa = np.empty(n)
    sli = np.empty(n, np.int64)
    S = L_.atsim_superlayers(alt.ctypes.data_as(C.POINTER(C.c_double)), cn2.ctypes.data_as(C.POINTER(C.c_double)),
                             n, bin_m, sa.ctypes.data_as(C.POINTER(C.c_double)),
                             sli.ctypes.data_as(C.POINTER(C.c_long)))
    if S < 0:
        raise AtsimError("atsim_superlayers: invalid profile (altitudes must be ascending)")
    return sa[:S].copy(), sli


class Simulation:
    """Multi-wavelength, multi-direction frame loop on one GPU.

    slayer_alt, sli: super-layer binning (see superlayers()); lambdas (W,);
    scoeff (W, L); unwrap_mode (W,) 0/1; xpos0, ypos0 (D, L) pixels;
    vx, vy (L,) pixels per second.
    """

    def __init__(self, ctx, N, Nout, M, slayer_alt, sli, lambdas, scoeff, unwrap_mode, xpos0, ypos0, vx, vy,
                 dx, dt, zangle=0.0, bicubic=True, fresnel=True, amplitude=True):
        self.ctx = ctx
        L = np.asarray(sli).size
        W = np.asarray(lambdas).size
        xpos0 = np.ascontiguousarray(np.asarray(xpos0, np.float64).reshape(-1, L))
        D = xpos0.shape[0]
        self._keep = dict(
            sa=np.ascontiguousarray(slayer_alt, np.float64), sli=np.ascontiguousarray(sli, np.int64),
            lam=np.ascontiguousarray(lambdas, np.float64),
            sc=np.ascontiguousarray(np.asarray(scoeff, np.float64).reshape(W, L)),
            uw=np.ascontiguousarray(unwrap_mode, np.int32), x0=xpos0,
            y0=np.ascontiguousarray(np.asarray(ypos0, np.float64).reshape(D, L)),
            vx=np.ascontiguousarray(vx, np.float64), vy=np.ascontiguousarray(vy, np.float64))
        k = self._keep
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        d = SimDesc()
        d.N, d.Nout, d.M, d.L, d.W, d.D = N, Nout, M, L, W, D
        d.dx, d.dt, d.zangle = dx, dt, zangle
        d.bicubic, d.fresnel, d.amplitude = int(bicubic), int(fresnel), int(amplitude)
        d.slayer_alt = dp(k["sa"])
        d.super_layer_index = k["sli"].ctypes.data_as(C.POINTER(C.c_long))
        d.lambda_ = dp(k["lam"])
        d.scoeff = dp(k["sc"])
        d.unwrap_mode = k["uw"].ctypes.data_as(C.POINTER(C.c_int))
        d.xpos0, d.ypos0, d.vx, d.vy = dp(k["x0"]), dp(k["y0"]), dp(k["vx"]), dp(k["vy"])
        h = C.c_void_p()
        ctx._chk(ctx.L.atsim_sim_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        ctx._sims.add(self)
        self.N, self.Nout, self.M, self.L, self.W, self.D = N, Nout, M, L, W, D
        torch = _torch()
        dev = f"cuda:{ctx.device}"
        self.pha = torch.empty((D, W, Nout, Nout), dtype=torch.float32, device=dev)
        self.amp = torch.empty_like(self.pha)

    def close(self):
        # a simulation outliving its context was already closed by it
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.ctx.L.atsim_sim_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def master_ptr(self, layer):
        return self.ctx.L.atsim_sim_master(self.h, layer)

    def set_master(self, layer, m):
        """m: numpy (M, M) array or torch float32 CUDA tensor"""
        torch = _torch()
        if isinstance(m, np.ndarray):
            a = np.ascontiguousarray(m, np.float32)
            assert a.size == self.M * self.M
            self.ctx._chk(self.ctx.L.atsim_sim_set_master(self.h, layer, a.ctypes.data_as(C.c_void_p)))
        else:
            assert m.dtype == torch.float32 and m.is_contiguous() and m.numel() == self.M * self.M
            torch.cuda.current_stream(m.device).synchronize()
            self.ctx._chk(self.ctx.L.atsim_sim_set_master(self.h, layer, _ptr(m)))

    def frame(self, field=None, sync=True, out=None):
        """advance one time step; returns (pha, amp) [D, W, Nout, Nout], in the
        simulation's own buffers or in out = (pha, amp) float32 device tensors"""
        pha, amp = out if out is not None else (self.pha, self.amp)
        if out is not None:
            assert pha.numel() == self.pha.numel() and amp.numel() == self.amp.numel()
            assert pha.dtype == amp.dtype == _torch().float32 and pha.is_contiguous() and amp.is_contiguous()
        self.ctx._chk(self.ctx.L.atsim_sim_frame(self.h, _ptr(pha), _ptr(amp), _ptr(field)))
        if sync:
            self.ctx.sync()
        return pha, amp

    def positions(self):
        x = np.empty((self.D, self.L))
        y = np.empty((self.D, self.L))
        self.ctx._chk(self.ctx.L.atsim_sim_positions(self.h, x.ctypes.data_as(C.POINTER(C.c_double)),
                                                     y.ctypes.data_as(C.POINTER(C.c_double))))
        return x, y

    def block_sums(self):
        """(D, W, xs, xs, 2) float64 block sums (sum cos, sum sin of PHI - pha)
        of the last frame's science-lambda rule, xs = Nout // 20"""
        xs = self.Nout // 20
        out = np.zeros((self.D, self.W, xs, xs, 2))
        self.ctx._chk(self.ctx.L.atsim_sim_block_sums(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), out.size))
        return out

    def autotune(self, frames=5, cache=None):
        """Pick the context's launch shapes for this simulation with the
        library's autotuner (atsim_sim_autotune) and return the winning
        settings.  cache: path of a JSON file of winners keyed by device and
        sizes (default $ATSIM_TUNING_CACHE, no file if unset); a cached entry
        is applied without a search."""
        import json
        import os
        torch = _torch()
        path = cache if cache is not None else os.environ.get("ATSIM_TUNING_CACHE")
        key = "|".join(str(v) for v in (torch.cuda.get_device_name(self.ctx.device), self.N, self.Nout, self.M,
                                        self.L, self.W, self.D))
        table = {}
        if path and os.path.exists(path):
            with open(path) as fh:
                table = json.load(fh)
        if key in table:
            for k, v in table[key].items():
                self.ctx.set_tuning(k, v)
            return dict(table[key], cached=True)
        buf = C.create_string_buffer(1 << 16)
        self.ctx._chk(self.ctx.L.atsim_sim_autotune(self.h, _ptr(self.pha), _ptr(self.amp), int(frames), buf,
                                                    len(buf)))
        lines = buf.value.decode().strip().splitlines()
        best = dict(kv.split("=") for kv in lines[-1].split()[1:])
        ms = float(best.pop("ms"))
        best = {k: int(v) for k, v in best.items()}
        if path:
            table[key] = best
            d = os.path.dirname(os.path.abspath(path))
            os.makedirs(d, exist_ok=True)
            with open(path, "w") as fh:
                json.dump(table, fh, indent=1)
        self.tune_report = lines
        return dict(best, ms_per_frame=ms, candidates=len(lines) - 1, cached=False)

    def use_masters(self, owner):
        """share the masters of another simulation of this context (same M and L)"""
        self.ctx._chk(self.ctx.L.atsim_sim_use_masters(self.h, owner.h))
        self._masters_owner = owner  # keep it alive

    def stats_enable(self, enable=True):
        """reduce every following frame's outputs to per-field statistics"""
        self.ctx._chk(self.ctx.L.atsim_sim_stats_enable(self.h, int(enable)))

    def stats(self):
        """(D, W, 4) float64 of the last frame: OPD piston [m], OPD rms [m],
        mean intensity, scintillation index"""
        out = np.zeros((self.D, self.W, 4))
        self.ctx._chk(self.ctx.L.atsim_sim_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), out.size))
        return out

    def gaps(self):
        return self.ctx.L.atsim_sim_gaps(self.h)

    def synth_masters(self, seed, outerscale_px, innerscale_px):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(outerscale_px, np.float64), (self.L,)))
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(innerscale_px, np.float64), (self.L,)))
        dp = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
        self.ctx._chk(self.ctx.L.atsim_sim_synth_masters(self.h, seed, dp(a), dp(b)))

    def normalize_masters(self, r0_goal_px, cn2, zangle=0.0):
        cn2 = np.ascontiguousarray(cn2, np.float64)
        r0 = np.empty(self.L)
        dp = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
        self.ctx._chk(self.ctx.L.atsim_sim_normalize_masters(self.h, r0_goal_px, dp(cn2), zangle, dp(r0)))
        return r0

    def master(self, layer):
        """torch float32 (M, M) copy of a layer's master"""
        torch = _torch()
        out = torch.empty((self.M, self.M), dtype=torch.float32, device=f"cuda:{self.ctx.device}")
        hip_copy(self.ctx, out.data_ptr(), self.master_ptr(layer), 4 * self.M * self.M)
        return out

    # row / col: passes with one Fresnel factor; row2 / col2: passes with a
    # factor, a layer group's screen and a factor (one pass per gap)
    PROF_KINDS = {"row": 0, "col": 1, "extract": 2, "block": 3, "row2": 4, "col2": 5, "phase": 6, "stats": 7,
                  "geometry": 8}

    def profile(self, enable=True):
        self.ctx._chk(self.ctx.L.atsim_sim_profile(self.h, int(enable)))

    def profile_read(self, kind):
        ms = C.c_double()
        n = C.c_long()
        self.ctx._chk(self.ctx.L.atsim_sim_profile_read(self.h, self.PROF_KINDS[kind], C.byref(ms), C.byref(n)))
        return ms.value, n.value


def sims_frame(sims, sync=True):
    """one frame of several simulations of one context (any grid sizes) in one
    call (atsim_sims_frame): each simulation's own (pha, amp) buffers"""
    ctx = sims[0].ctx
    n = len(sims)
    hs = (C.c_void_p * n)(*[s.h.value if isinstance(s.h, C.c_void_p) else s.h for s in sims])
    ph = (C.c_void_p * n)(*[s.pha.data_ptr() for s in sims])
    am = (C.c_void_p * n)(*[s.amp.data_ptr() for s in sims])
    ctx._chk(ctx.L.atsim_sims_frame(hs, n, ph, am))
    if sync:
        ctx.sync()
    return [(s.pha, s.amp) for s in sims]
