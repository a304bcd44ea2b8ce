"""Fourier-neural-operator surrogate of the layer loop (no counterpart in the
reference): a per-layer phase stack -> pupil log-amplitude and diffraction
phase, trained on pairs the simulator generates, served with the plain
Fresnel path as judge and fallback.

Every operation runs in the library's own kernels through the C ABI
(atsim_gemm_bf16 on the matrix cores, atsim_nn_*); torch only holds device
memory, on the library context's stream.

Network and conventions (the same as the FP64 restatement used to check it):
    h0 = Lw x + Lb;  h_{b+1} = act(Wb h_b + Bb + K_b h_b);  y = Q2 GELU(Q1 h + q1b) + q2b
    K_b h = Re IDFT2(R_b(k) DFT2(h)(k)) over ky in [-m, m), kx in [0, m),
    kx > 0 doubled in the inverse (real-input convention), 1/P scaling.
Activations are channel-major over the batch: [C][S][N][N]."""
import ctypes as C
import math

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def on_ctx_stream(fn):
    """run a method with torch's current stream set to the library context's
    stream, so torch allocations and copies are ordered with the kernels"""
    def wrapped(self, *a, **k):
        torch = _torch()
        ctx = self.ctx if hasattr(self, "ctx") else (a[0] if a else k["ctx"])
        with torch.cuda.stream(ctx.stream):
            return fn(self, *a, **k)
    wrapped.__name__ = fn.__name__
    wrapped.__doc__ = fn.__doc__
    return wrapped


class Ops:
    """Thin wrappers of the network entry points of one context."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.L = ctx.L

    def gemm(self, M, N, K, batch, A, a_off, sab, sam, sak, B, b_off, sbb, sbk, sbn, Cm, c_off, scb, scm, scn,
             alpha=1.0, beta=0.0, bias=None, pre=None, gelu=False):
        """C = act(alpha A B + beta C + bias[m]) on float32 tensors with element
        offsets and strides (pre, if given, receives the value before the
        activation); raises before the launch if any operand's last element
        lies outside its tensor"""
        if M > 0 and N > 0 and batch > 0:
            last = lambda off, sb, d: off + (batch - 1) * sb + sum((n - 1) * st for n, st in d)
            if (min(a_off, b_off, c_off) < 0
                    or last(a_off, sab, [(M, sam), (K, sak)]) >= A.numel()
                    or last(b_off, sbb, [(K, sbk), (N, sbn)]) >= B.numel()
                    or last(c_off, scb, [(M, scm), (N, scn)]) >= Cm.numel()
                    or (bias is not None and bias.numel() < M)
                    or (pre is not None and pre.numel() != Cm.numel())):
                raise ValueError(f"gemm operands out of bounds (M {M} N {N} K {K} batch {batch})")
        L = self.L
        pa = C.c_void_p(A.data_ptr() + 4 * a_off)
        pb = C.c_void_p(B.data_ptr() + 4 * b_off)
        pc = C.c_void_p(Cm.data_ptr() + 4 * c_off)
        if bias is None and pre is None and not gelu:
            self.ctx._chk(L.atsim_gemm_bf16(self.ctx.h, M, N, K, batch, alpha, pa, sab, sam, sak, pb, sbb, sbk, sbn,
                                            beta, pc, scb, scm, scn))
        else:
            pp = C.c_void_p(pre.data_ptr() + 4 * c_off) if pre is not None else None
            self.ctx._chk(L.atsim_gemm_bf16_ex(self.ctx.h, M, N, K, batch, alpha, pa, sab, sam, sak, pb, sbb, sbk,
                                               sbn, beta, pc, scb, scm, scn, _p(bias), pp, int(gelu)))

    def add_bias_act(self, a, b, bias, pre, y, Cc, P, gelu):
        n = a.numel()
        if any(t is not None and t.numel() != n for t in (b, pre, y)) or (bias is not None and bias.numel() != Cc):
            raise ValueError("add_bias_act: operand sizes differ")
        self.ctx._chk(self.L.atsim_nn_add_bias_act(self.ctx.h, _p(a), _p(b), _p(bias), _p(pre), _p(y), Cc, P,
                                                   a.numel(), int(gelu)))

    def gelu_bwd(self, pre, dy, dx):
        if not pre.numel() == dy.numel() == dx.numel():
            raise ValueError(f"gelu_bwd: sizes {pre.numel()}, {dy.numel()}, {dx.numel()} differ")
        self.ctx._chk(self.L.atsim_nn_gelu_bwd(self.ctx.h, _p(pre), _p(dy), _p(dx), pre.numel()))

    def bias_grad(self, dy, db, Cc, P):
        if dy.numel() != Cc * P or db.numel() != Cc:
            raise ValueError("bias_grad: operand sizes differ")
        self.ctx._chk(self.L.atsim_nn_bias_grad(self.ctx.h, _p(dy), _p(db), 1, Cc, P, 0))

    def mse(self, pred, target, grad, want_loss=True):
        if target.numel() != pred.numel() or (grad is not None and grad.numel() != pred.numel()):
            raise ValueError("mse: operand sizes differ")
        loss = C.c_double()
        self.ctx._chk(self.L.atsim_nn_mse(self.ctx.h, _p(pred), _p(target), _p(grad), pred.numel(),
                                          C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def reduce(self, part, out, nb, n):
        if part.numel() < nb * n or out.numel() < n:
            raise ValueError("reduce: operands too small")
        self.ctx._chk(self.L.atsim_nn_reduce_batches(self.ctx.h, _p(part), _p(out), nb, n, 0))

    def adam(self, w, g, m, v, lr, b1, b2, eps, step):
        if not w.numel() == g.numel() == m.numel() == v.numel():
            raise ValueError("adam: operand sizes differ")
        self.ctx._chk(self.L.atsim_nn_adam(self.ctx.h, _p(w), _p(g), _p(m), _p(v), w.numel(), lr, b1, b2, eps,
                                           step))


def init_params(cin, cout, width, modes, blocks, hidden, seed=0):
    """float64 numpy parameters; R_b complex [2m][m][width][width]"""
    rng = np.random.default_rng(seed)
    u = lambda shape, fan: rng.uniform(-1.0, 1.0, shape) / math.sqrt(fan)
    p = {"Lw": u((width, cin), cin), "Lb": u((width,), cin)}
    for b in range(blocks):
        p[f"W{b}"] = u((width, width), width)
        p[f"B{b}"] = u((width,), width)
        sc = 1.0 / (width * width)
        p[f"R{b}"] = sc * (rng.uniform(0, 1, (2 * modes, modes, width, width))
                           + 1j * rng.uniform(0, 1, (2 * modes, modes, width, width)))
    p["Q1"] = u((hidden, width), width)
    p["q1b"] = u((hidden,), width)
    p["Q2"] = u((cout, hidden), hidden)
    p["q2b"] = u((cout,), hidden)
    return p


class FNO:
    """The network on the GPU: parameters, forward, backward, Adam."""

    @on_ctx_stream
    def __init__(self, ctx, N, cin, cout, width=32, modes=16, blocks=4, hidden=64, params=None, seed=0,
                 device=None):
        torch = _torch()
        self.ctx, self.ops = ctx, Ops(ctx)
        self.N, self.cin, self.cout, self.width, self.m, self.blocks, self.hidden = N, cin, cout, width, modes, \
            blocks, hidden
        self.dev = device or f"cuda:{ctx.device}"
        if params is None:
            params = init_params(cin, cout, width, modes, blocks, hidden, seed)
        self.param = {}
        for k, v in params.items():
            a = np.asarray(v)
            if np.iscomplexobj(a):
                a = np.stack([a.real, a.imag], axis=-1)
            self.param[k] = torch.tensor(np.ascontiguousarray(a, np.float32), device=self.dev)
        self.grad = {k: torch.zeros_like(v) for k, v in self.param.items()}
        self.adam_m = {k: torch.zeros_like(v) for k, v in self.param.items()}
        self.adam_v = {k: torch.zeros_like(v) for k, v in self.param.items()}
        self.step = 0
        self._tables()

    # ---------------------------------------------------------------- DFT tables
    def _tables(self):  # called from __init__ (on the context stream)
        torch = _torch()
        N, m = self.N, self.m
        x = np.arange(N)
        kx = np.arange(m)
        ky = np.concatenate([np.arange(m), np.arange(N - m, N)])
        ax = 2 * np.pi * np.outer(x, kx) / N                   # [N][m]
        Fx = np.empty((N, 2 * m))
        Fx[:, 0::2] = np.cos(ax)
        Fx[:, 1::2] = -np.sin(ax)
        ay = 2 * np.pi * np.outer(ky, x) / N                   # [2m][N]
        w = np.where(kx == 0, 1.0, 2.0) / (N * N)
        Fxi = np.empty((2 * m, N))
        Fxi[0::2, :] = (w[:, None] * np.cos(ax.T))
        Fxi[1::2, :] = (-w[:, None] * np.sin(ax.T))
        f = lambda a: torch.tensor(np.ascontiguousarray(a, np.float32), device=self.dev)
        self.Fx, self.Gre, self.Gim, self.Fxi = f(Fx), f(np.cos(ay)), f(-np.sin(ay)), f(Fxi)

    # ---------------------------------------------------------------- building blocks
    def _pointwise(self, Wt, h, out, Cout, Cin, SP, beta=0.0, bias=None, pre=None, gelu=False):
        # out[o][q] = act(sum_i W[o][i] h[i][q] + beta out[o][q] + bias[o])
        self.ops.gemm(Cout, SP, Cin, 1, Wt, 0, 0, Cin, 1, h, 0, 0, SP, 1, out, 0, 0, SP, 1, beta=beta, bias=bias,
                      pre=pre, gelu=gelu)

    def _pointwise_t(self, Wt, dy, out, Cout, Cin, SP, beta=0.0):
        # out[i][q] = sum_o W[o][i] dy[o][q]
        self.ops.gemm(Cin, SP, Cout, 1, Wt, 0, 0, 1, Cin, dy, 0, 0, SP, 1, out, 0, 0, SP, 1, beta=beta)

    def _weight_grad(self, dy, h, gW, Cout, Cin, SP):
        # gW[o][i] = sum_q dy[o][q] h[i][q], split over K chunks, partials summed in order
        torch = _torch()
        kc = 1024 if SP % 1024 == 0 else SP
        nb = SP // kc
        part = torch.empty((nb, Cout, Cin), dtype=torch.float32, device=self.dev)
        self.ops.gemm(Cout, Cin, kc, nb, dy, 0, kc, SP, 1, h, 0, kc, 1, SP, part, 0, Cout * Cin, Cin, 1)
        self.ops.reduce(part, gW, nb, Cout * Cin)

    def _spectral(self, h, R, S, out):
        """out [C][S][N][N] = K h; returns the retained spectrum B [C][S][2m][m][2]"""
        torch = _torch()
        N, m, Cw = self.N, self.m, self.width
        rows = Cw * S * N
        cs = Cw * S
        T1 = torch.empty((rows, 2 * m), dtype=torch.float32, device=self.dev)
        self.ops.gemm(rows, 2 * m, N, 1, h, 0, 0, N, 1, self.Fx, 0, 0, 2 * m, 1, T1, 0, 0, 2 * m, 1)
        B = torch.empty((Cw, S, 2 * m, m, 2), dtype=torch.float32, device=self.dev)
        self._cmat_left(self.Gre, self.Gim, T1, B, cs, transpose=False)
        O = torch.empty_like(B)
        self._mix(R, B, O, S)
        U = torch.empty((rows, 2 * m), dtype=torch.float32, device=self.dev)
        self._cmat_left_inv(O, U, cs)
        self.ops.gemm(rows, N, 2 * m, 1, U, 0, 0, 2 * m, 1, self.Fxi, 0, 0, N, 1, out, 0, 0, N, 1)
        return B

    def _cmat_left(self, Gre, Gim, T, Bout, batch, transpose):
        """Bout[b] (2m x m complex) = G (2m x N complex) T[b] (N x m complex, interleaved columns)"""
        N, m = self.N, self.m
        g = self.ops.gemm
        ts, bs = N * 2 * m, 4 * m * m
        # Bre = Gre Tre - Gim Tim ; Bim = Gre Tim + Gim Tre
        g(2 * m, m, N, batch, Gre, 0, 0, N, 1, T, 0, ts, 2 * m, 2, Bout, 0, bs, 2 * m, 2)
        g(2 * m, m, N, batch, Gim, 0, 0, N, 1, T, 1, ts, 2 * m, 2, Bout, 0, bs, 2 * m, 2, alpha=-1.0, beta=1.0)
        g(2 * m, m, N, batch, Gre, 0, 0, N, 1, T, 1, ts, 2 * m, 2, Bout, 1, bs, 2 * m, 2)
        g(2 * m, m, N, batch, Gim, 0, 0, N, 1, T, 0, ts, 2 * m, 2, Bout, 1, bs, 2 * m, 2, beta=1.0)

    def _cmat_left_inv(self, O, U, batch):
        """U[b] (N x m complex, interleaved) = G^H O[b]: Ure = Gre^T Ore + Gim^T Oim, Uim = Gre^T Oim - Gim^T Ore"""
        N, m = self.N, self.m
        g = self.ops.gemm
        us, os_ = N * 2 * m, 4 * m * m
        g(N, m, 2 * m, batch, self.Gre, 0, 0, 1, N, O, 0, os_, 2 * m, 2, U, 0, us, 2 * m, 2)
        g(N, m, 2 * m, batch, self.Gim, 0, 0, 1, N, O, 1, os_, 2 * m, 2, U, 0, us, 2 * m, 2, beta=1.0)
        g(N, m, 2 * m, batch, self.Gre, 0, 0, 1, N, O, 1, os_, 2 * m, 2, U, 1, us, 2 * m, 2)
        g(N, m, 2 * m, batch, self.Gim, 0, 0, 1, N, O, 0, os_, 2 * m, 2, U, 1, us, 2 * m, 2, alpha=-1.0, beta=1.0)

    def _mix(self, R, B, O, S):
        """O[o][s][k] = sum_i R[k][o][i] B[i][s][k] (complex), one GEMM batch per mode"""
        m, Cw = self.m, self.width
        nm = 2 * m * m
        g = self.ops.gemm
        ra = (2 * Cw * Cw, 2 * Cw, 2)          # sab, sam, sak of R
        bb = (2, S * 2 * nm, 2 * nm)           # sbb, sbk, sbn of B (i, s)
        cc = (2, S * 2 * nm, 2 * nm)
        g(Cw, S, Cw, nm, R, 0, *ra, B, 0, *bb, O, 0, *cc)
        g(Cw, S, Cw, nm, R, 1, *ra, B, 1, *bb, O, 0, *cc, alpha=-1.0, beta=1.0)
        g(Cw, S, Cw, nm, R, 0, *ra, B, 1, *bb, O, 1, *cc)
        g(Cw, S, Cw, nm, R, 1, *ra, B, 0, *bb, O, 1, *cc, beta=1.0)

    def _spectral_backward(self, g_out, R, B, S, dh, dR):
        """dh (written) and dR (written) from the upstream gradient g_out [C][S][N][N]"""
        torch = _torch()
        N, m, Cw = self.N, self.m, self.width
        rows, cs, nm = Cw * S * N, Cw * S, 2 * m * m
        gm = self.ops.gemm
        dU = torch.empty((rows, 2 * m), dtype=torch.float32, device=self.dev)
        gm(rows, 2 * m, N, 1, g_out, 0, 0, N, 1, self.Fxi, 0, 0, 1, N, dU, 0, 0, 2 * m, 1)
        dO = torch.empty((Cw, S, 2 * m, m, 2), dtype=torch.float32, device=self.dev)
        # dOre = Gre dUre - Gim dUim ; dOim = Gim dUre + Gre dUim
        self._cmat_left(self.Gre, self.Gim, dU, dO, cs, transpose=False)
        # (the same form as the forward y-DFT: G applied to the interleaved dU)
        dB = torch.empty_like(dO)
        ra_t = (2 * Cw * Cw, 2, 2 * Cw)        # R^T: (i, o) -> sam = 2, sak = 2 Cw
        ob = (2, S * 2 * nm, 2 * nm)
        # dBre = Rre^T dOre + Rim^T dOim ; dBim = Rre^T dOim - Rim^T dOre
        gm(Cw, S, Cw, nm, R, 0, *ra_t, dO, 0, *ob, dB, 0, *ob)
        gm(Cw, S, Cw, nm, R, 1, *ra_t, dO, 1, *ob, dB, 0, *ob, beta=1.0)
        gm(Cw, S, Cw, nm, R, 0, *ra_t, dO, 1, *ob, dB, 1, *ob)
        gm(Cw, S, Cw, nm, R, 1, *ra_t, dO, 0, *ob, dB, 1, *ob, alpha=-1.0, beta=1.0)
        # dRre = dOre Bre^T + dOim Bim^T ; dRim = dOim Bre^T - dOre Bim^T   (contract over s)
        oa = (2, S * 2 * nm, 2 * nm)           # dO as (o, s): sab, sam, sak
        bt = (2, 2 * nm, S * 2 * nm)           # B^T as (s, i): sbb, sbk, sbn
        rc = (2 * Cw * Cw, 2 * Cw, 2)
        gm(Cw, Cw, S, nm, dO, 0, *oa, B, 0, *bt, dR, 0, *rc)
        gm(Cw, Cw, S, nm, dO, 1, *oa, B, 1, *bt, dR, 0, *rc, beta=1.0)
        gm(Cw, Cw, S, nm, dO, 1, *oa, B, 0, *bt, dR, 1, *rc)
        gm(Cw, Cw, S, nm, dO, 0, *oa, B, 1, *bt, dR, 1, *rc, alpha=-1.0, beta=1.0)
        # dT1re = Gre^T dBre + Gim^T dBim ; dT1im = Gre^T dBim - Gim^T dBre  (the inverse-y form)
        dT1 = torch.empty((rows, 2 * m), dtype=torch.float32, device=self.dev)
        self._cmat_left_inv(dB, dT1, cs)
        # dh = dT1 Fx^T
        gm(rows, N, 2 * m, 1, dT1, 0, 0, 2 * m, 1, self.Fx, 0, 0, 1, 2 * m, dh, 0, 0, N, 1)

    # ---------------------------------------------------------------- forward / backward
    @on_ctx_stream
    def forward(self, x, keep=True):
        """x [cin][S][N][N] float32 -> y [cout][S][N][N]"""
        torch = _torch()
        P = self.N * self.N
        S = x.shape[1]
        SP = S * P
        Cw, p, ops = self.width, self.param, self.ops
        cache = {"x": x, "S": S}
        h = torch.empty((Cw, S, self.N, self.N), dtype=torch.float32, device=self.dev)
        self._pointwise(p["Lw"], x, h, Cw, self.cin, SP, bias=p["Lb"])
        for b in range(self.blocks):
            # hn = act(W h + K h + B): the spectral path writes K h, the 1x1 map
            # adds onto it with its bias and activation in one GEMM epilogue
            hn = torch.empty_like(h)
            B = self._spectral(h, p[f"R{b}"], S, hn)
            pre = torch.empty_like(h) if keep else None
            self._pointwise(p[f"W{b}"], h, hn, Cw, Cw, SP, beta=1.0, bias=p[f"B{b}"], pre=pre,
                            gelu=b < self.blocks - 1)
            if keep:
                cache[f"h{b}"], cache[f"pre{b}"], cache[f"B{b}"] = h, pre, B
            h = hn
        cache["hL"] = h
        q1 = torch.empty((self.hidden, S, self.N, self.N), dtype=torch.float32, device=self.dev)
        preq = torch.empty_like(q1) if keep else None
        self._pointwise(p["Q1"], h, q1, self.hidden, Cw, SP, bias=p["q1b"], pre=preq, gelu=True)
        y = torch.empty((self.cout, S, self.N, self.N), dtype=torch.float32, device=self.dev)
        self._pointwise(p["Q2"], q1, y, self.cout, self.hidden, SP, bias=p["q2b"])
        if keep:
            cache["preq"], cache["q1"] = preq, q1
            self.cache = cache
        return y

    @on_ctx_stream
    def loss_and_grads(self, x, t, want_loss=True):
        """forward + backward; gradients land in self.grad; returns the MSE loss
        (None without want_loss, which spares a stream synchronisation)"""
        torch = _torch()
        y = self.forward(x, keep=True)
        c, p, g, ops = self.cache, self.param, self.grad, self.ops
        S = c["S"]
        P = self.N * self.N
        SP = S * P
        Cw, H = self.width, self.hidden
        dy = torch.empty_like(y)
        loss = ops.mse(y, t, dy, want_loss)
        self._weight_grad(dy, c["q1"], g["Q2"], self.cout, H, SP)
        ops.bias_grad(dy, g["q2b"], self.cout, SP)
        dq1 = torch.empty_like(c["q1"])
        self._pointwise_t(p["Q2"], dy, dq1, self.cout, H, SP)
        ops.gelu_bwd(c["preq"], dq1, dq1)
        self._weight_grad(dq1, c["hL"], g["Q1"], H, Cw, SP)
        ops.bias_grad(dq1, g["q1b"], H, SP)
        dh = torch.empty_like(c["hL"])
        self._pointwise_t(p["Q1"], dq1, dh, H, Cw, SP)
        for b in reversed(range(self.blocks)):
            if b < self.blocks - 1:
                ops.gelu_bwd(c[f"pre{b}"], dh, dh)
            dpre = dh
            h = c[f"h{b}"]
            ops.bias_grad(dpre, g[f"B{b}"], Cw, SP)
            self._weight_grad(dpre, h, g[f"W{b}"], Cw, Cw, SP)
            dhn = torch.empty_like(h)
            self._spectral_backward(dpre, p[f"R{b}"], c[f"B{b}"], S, dhn, g[f"R{b}"])
            self._pointwise_t(p[f"W{b}"], dpre, dhn, Cw, Cw, SP, beta=1.0)
            dh = dhn
        self._weight_grad(dh, c["x"], g["Lw"], Cw, self.cin, SP)
        ops.bias_grad(dh, g["Lb"], Cw, SP)
        return loss

    @on_ctx_stream
    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        for k in self.param:
            self.ops.adam(self.param[k], self.grad[k], self.adam_m[k], self.adam_v[k], lr, b1, b2, eps, self.step)

    @on_ctx_stream
    def params_numpy(self):
        out = {}
        for k, v in self.param.items():
            a = v.cpu().numpy().astype(np.float64)
            if k.startswith("R"):
                a = a[..., 0] + 1j * a[..., 1]
            out[k] = a
        return out

    @on_ctx_stream
    def grads_numpy(self):
        out = {}
        for k, v in self.grad.items():
            a = v.cpu().numpy().astype(np.float64)
            if k.startswith("R"):
                a = a[..., 0] + 1j * a[..., 1]
            out[k] = a
        return out
