"""ctypes view of the CPU oracle (oracle/libatsim_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.  Complex arrays
cross the boundary as numpy complex128 (interleaved re, im doubles), real
arrays as float64, row-major with x fastest (shape (n1, n0)).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libatsim_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)


def build(force=False):
    """Compile the oracle with its Makefile (gcc)."""
    src = os.path.join(_HERE, "atsim_oracle.c")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_fft1d.argtypes = [_dp, C.c_long, C.c_int]
        L.orc_fft2d.argtypes = [_dp, C.c_long, C.c_long, C.c_int]
        L.orc_dft1d_naive.argtypes = [_dp, _dp, C.c_long, C.c_int]
        L.orc_permut_c.argtypes = [_dp, C.c_long, C.c_long]
        L.orc_permut_r.argtypes = [_dp, C.c_long, C.c_long]
        L.orc_fresnel.argtypes = [_dp, _dp, C.c_long, C.c_long, C.c_double, C.c_double, C.c_double]
        L.orc_fresnel_init_tf.argtypes = [_dp, C.c_long, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_fresnel1.argtypes = [_dp, _dp, _dp, C.c_long, C.c_long]
        L.orc_fresnel_cube.argtypes = [_dp, _dp, _dp, C.c_long, C.c_long, C.c_double, C.c_double,
                                       C.c_double, C.c_long, C.c_double]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_noise_c64.argtypes = [C.c_uint64, C.c_uint32, C.c_long, _dp, C.POINTER(C.c_float)]
        L.orc_vk_filter.argtypes = [_dp, C.c_long, C.c_double, C.c_double, C.c_double]
        L.orc_strfn_direct.argtypes = [_dp, C.c_long, C.c_long, C.c_long]
        L.orc_strfn_direct.restype = C.c_double
        L.orc_strfn_fft.argtypes = [_dp, C.c_long, _dp]
        L.orc_strfn_coeff.argtypes = [_dp, C.c_long]
        L.orc_strfn_coeff.restype = C.c_double
        L.orc_make_master.argtypes = [_dp, C.c_long, C.c_double, C.c_double, C.c_double, C.c_int,
                                      _dp, _dp, _dp, _dp]
        L.orc_r0_measure.argtypes = [_dp, C.c_long, C.c_long]
        L.orc_r0_measure.restype = C.c_double
        L.orc_superlayers.argtypes = [_dp, _dp, C.c_long, C.c_double, _dp, _lp]
        L.orc_superlayers.restype = C.c_long
        L.orc_advance_position.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, C.c_long]
        L.orc_sample.argtypes = [_dp, C.c_long, C.c_double, C.c_double, C.c_long, C.c_long, C.c_int]
        L.orc_sample.restype = C.c_double
        L.orc_sim_frame.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_extract_simple.argtypes = [_dp, _dp, C.c_long, C.c_long, _dp, _dp]
        L.orc_extract_block.argtypes = [_dp, _dp, C.c_long, C.c_long, C.c_int, _dp, _dp, _dp]
        L.orc_extract_block.restype = C.c_long
        L.orc_block_fix.argtypes = [_dp, C.c_long, C.c_long, C.c_int, _dp, C.c_long]
        L.orc_block_fix.restype = C.c_long
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- FFT
def fft1d(x, sign=-1):
    y = _c128(x).copy()
    assert lib().orc_fft1d(_p(y), y.size, sign) == 0
    return y


def dft1d_naive(x, sign=-1):
    x = _c128(x)
    y = np.empty_like(x)
    lib().orc_dft1d_naive(_p(x), _p(y), x.size, sign)
    return y


def fft2d(x, sign=-1):
    y = _c128(x).copy()
    n1, n0 = y.shape
    assert lib().orc_fft2d(_p(y), n0, n1, sign) == 0
    return y


def permut(x):
    if np.iscomplexobj(x):
        y = _c128(x).copy()
        n1, n0 = y.shape
        lib().orc_permut_c(_p(y), n0, n1)
    else:
        y = _f64(x).copy()
        n1, n0 = y.shape
        lib().orc_permut_r(_p(y), n0, n1)
    return y


# ---------------------------------------------------------------- Fresnel
def fresnel(E, dx, z, lam):
    E = _c128(E)
    out = np.empty_like(E)
    n1, n0 = E.shape
    assert lib().orc_fresnel(_p(E), _p(out), n0, n1, dx, z, lam) == 0
    return out


def fresnel_init_tf(size, dx, z, lam, fpmaskrad):
    C_ = np.empty((size, size), np.complex128)
    lib().orc_fresnel_init_tf(_p(C_), size, dx, z, lam, fpmaskrad)
    return C_


def fresnel1(E, Ctf):
    E = _c128(E)
    Ctf = _c128(Ctf)
    out = np.empty_like(E)
    n1, n0 = E.shape
    assert lib().orc_fresnel1(_p(E), _p(out), _p(Ctf), n0, n1) == 0
    return out


def fresnel_cube(E, dx, zstart, zend, nz, lam):
    E = _c128(E)
    n1, n0 = E.shape
    amp = np.empty((nz, n1, n0))
    pha = np.empty((nz, n1, n0))
    assert lib().orc_fresnel_cube(_p(E), _p(amp), _p(pha), n0, n1, dx, zstart, zend, nz, lam) == 0
    return amp, pha


# ---------------------------------------------------------------- noise
def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def noise_c64(seed, stream, n, f32=False):
    out64 = np.empty(n, np.complex128)
    if f32:
        out32 = np.empty(2 * n, np.float32)
        lib().orc_noise_c64(seed, stream, n, _p(out64), out32.ctypes.data_as(C.POINTER(C.c_float)))
        return out64, out32.view(np.complex64)
    lib().orc_noise_c64(seed, stream, n, _p(out64), None)
    return out64


# ---------------------------------------------------------------- synthesis
def vk_filter(size, L0px, l0px, rlim=0.0):
    A = np.empty((size, size))
    lib().orc_vk_filter(_p(A), size, L0px, l0px, rlim)
    return A


def strfn_direct(s, a, b):
    s = _f64(s)
    return lib().orc_strfn_direct(_p(s), s.shape[0], a, b)


def strfn_fft(s):
    s = _f64(s)
    D = np.empty_like(s)
    lib().orc_strfn_fft(_p(s), s.shape[0], _p(D))
    return D


def strfn_coeff(s):
    s = _f64(s)
    return lib().orc_strfn_coeff(_p(s), s.shape[0])


def make_master(noise, L0px, l0px, rlim=0.0, prec_f32=False):
    noise = _c128(noise)
    size = noise.shape[0]
    o1 = np.empty((size, size))
    o2 = np.empty((size, size))
    c1 = C.c_double()
    c2 = C.c_double()
    assert lib().orc_make_master(_p(noise), size, L0px, l0px, rlim, int(prec_f32), _p(o1), _p(o2),
                                 C.pointer(c1), C.pointer(c2)) == 0
    return o1, o2, c1.value, c2.value


def r0_measure(m, dpix=50):
    m = _f64(m)
    return lib().orc_r0_measure(_p(m), m.shape[0], dpix)


def superlayers(alt, cn2, bin_m):
    alt = _f64(alt)
    cn2 = _f64(cn2)
    L = alt.size
    sa = np.empty(L)
    sli = np.empty(L, dtype=np.int64)
    S = lib().orc_superlayers(_p(alt), _p(cn2), L, bin_m, _p(sa), sli.ctypes.data_as(_lp))
    return sa[:S].copy(), sli


def advance_position(x, y, vx, vy, dt, M):
    xx = C.c_double(x)
    yy = C.c_double(y)
    lib().orc_advance_position(C.byref(xx), C.byref(yy), vx, vy, dt, M)
    return xx.value, yy.value


def sample(master, xpos, ypos, ii, jj, bicubic=True):
    m = _f64(master)
    return lib().orc_sample(_p(m), m.shape[0], xpos, ypos, ii, jj, int(bicubic))


class _Sim(C.Structure):
    _fields_ = [
        ("N", C.c_long), ("Nout", C.c_long), ("M", C.c_long), ("L", C.c_long), ("W", C.c_long),
        ("dx", C.c_double), ("dt", C.c_double), ("zangle", C.c_double),
        ("bicubic", C.c_int), ("fresnel", C.c_int), ("amplitude", C.c_int), ("prec_f32", C.c_int),
        ("masters", _dp), ("slayer_alt", _dp), ("super_layer_index", _lp),
        ("lambda_", _dp), ("scoeff", _dp), ("unwrap_mode", C.POINTER(C.c_int)),
        ("xpos", _dp), ("ypos", _dp), ("vx", _dp), ("vy", _dp),
    ]


class Sim:
    """Frame loop of the reference (AtmosphericTurbulence.c:2561-4262) over W fields.

    masters: (L, M, M) float64, already normalised/weighted per layer.
    scoeff:  (W, L) phase multipliers; unwrap_mode: (W,) 0 = reference-lambda
    rule, 1 = science-lambda block rule.
    """

    def __init__(self, N, Nout, masters, slayer_alt, sli, lambdas, scoeff, unwrap_mode,
                 xpos, ypos, vx, vy, dx, dt, zangle=0.0, bicubic=True, fresnel=True,
                 amplitude=True, prec_f32=False):
        self.masters = _f64(masters)
        L, M, _ = self.masters.shape
        self.slayer_alt = _f64(slayer_alt)
        self.sli = np.ascontiguousarray(sli, dtype=np.int64)
        self.lambdas = _f64(lambdas)
        W = self.lambdas.size
        self.scoeff = _f64(np.asarray(scoeff, dtype=np.float64).reshape(W, L))
        self.unwrap = np.ascontiguousarray(unwrap_mode, dtype=np.int32)
        self.xpos = _f64(xpos).copy()
        self.ypos = _f64(ypos).copy()
        self.vx = _f64(vx)
        self.vy = _f64(vy)
        self.N, self.Nout, self.L, self.W, self.M = N, Nout, L, W, M
        s = _Sim()
        s.N, s.Nout, s.M, s.L, s.W = N, Nout, M, L, W
        s.dx, s.dt, s.zangle = dx, dt, zangle
        s.bicubic, s.fresnel, s.amplitude, s.prec_f32 = int(bicubic), int(fresnel), int(amplitude), int(prec_f32)
        s.masters = _p(self.masters)
        s.slayer_alt = _p(self.slayer_alt)
        s.super_layer_index = self.sli.ctypes.data_as(_lp)
        s.lambda_ = _p(self.lambdas)
        s.scoeff = _p(self.scoeff)
        s.unwrap_mode = self.unwrap.ctypes.data_as(C.POINTER(C.c_int))
        s.xpos = _p(self.xpos)
        s.ypos = _p(self.ypos)
        s.vx = _p(self.vx)
        s.vy = _p(self.vy)
        self._s = s

    def frame(self, want_field=True):
        N, Nout, W = self.N, self.Nout, self.W
        E = np.empty((W, N, N), np.complex128) if want_field else None
        PHI = np.empty((W, N, N)) if want_field else None
        pha = np.empty((W, Nout, Nout))
        amp = np.empty((W, Nout, Nout))
        rc = lib().orc_sim_frame(C.byref(self._s), _p(E) if E is not None else None,
                                 _p(PHI) if PHI is not None else None, _p(pha), _p(amp))
        assert rc == 0
        return E, PHI, pha, amp


def extract_simple(E, PHI, Nout):
    E = _c128(E)
    PHI = _f64(PHI)
    N = E.shape[0]
    pha = np.empty((Nout, Nout))
    amp = np.empty((Nout, Nout))
    lib().orc_extract_simple(_p(E), _p(PHI), N, Nout, _p(pha), _p(amp))
    return pha, amp


def extract_block(E, PHI, Nout, prec_f32=False):
    E = _c128(E)
    PHI = _f64(PHI)
    N = E.shape[0]
    pha = np.empty((Nout, Nout))
    amp = np.empty((Nout, Nout))
    nb = (Nout // 20) ** 2
    peak = np.empty(max(nb, 1))
    it = lib().orc_extract_block(_p(E), _p(PHI), N, Nout, int(prec_f32), _p(pha), _p(amp), _p(peak))
    return pha, amp, peak[:nb].reshape(Nout // 20, Nout // 20), it


def block_fix(bin_, prec_f32=False, max_rounds=1000):
    """The block fix-up loop on a (ys, xs) grid of peaks (AtmosphericTurbulence.c:3627-4216).

    max_rounds <= 1000 stops the loop early (the reference's cap is 1000).
    Returns (peaks after the loop, first-round counts, number of rounds)."""
    b = np.array(bin_, dtype=np.float64, order="C", copy=True)
    ys, xs = b.shape
    ch = np.zeros_like(b)
    it = lib().orc_block_fix(_p(b), xs, ys, int(prec_f32), _p(ch), int(max_rounds))
    return b, ch, it
