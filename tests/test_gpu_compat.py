"""GPU parity of the reference-compatible entry points (include/atsim_compat.h)
against the oracle.

* `atsim mkwfs` (make_AtmosphericTurbulence_wavefront_series,
  AtmosphericTurbulence.c:1433-4455) end to end in a scratch directory: the
  masters it writes (t%03ld_%ld_f.fits) are normalised as the reference does
  (r0 per master, (r0 / r0goal)^(5/6), sqrt(cn2 / cos z)), the oracle runs the
  same frames, and the FITS cubes at the reference and science wavelengths,
  the science shared-memory image and SKIP_EXISTING are checked.
* make_master_turbulence_screen, do2dfft / do2dffti, permut,
  Fresnel_propagate_wavefront and Fresnel_propagate_cube on registry images.

Bar: relative L2 1e-5 for fields, amplitudes, FFTs and screens; phase within
relative L2 1e-5 on pixels whose 2 pi decision agrees, at most 0.1 % of pixels
disagreeing; permutation bit-exact."""
import ctypes as C
import math
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from tests import fitsutil
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "atmosphericturbulencesimul_amd", "atsim")
f32 = lambda v: float(np.float32(v))  # noqa: E731  (configuration scalars are float)


def _n0(lam):
    s2 = 1.0 / (lam * 1e6) ** 2
    return 0.0000834213 + 0.0240603 / (130.0 - s2) + 0.00015997 / (38.9 - s2)


def _roffset(site, alt, z, lam):
    sz = math.sin(z)
    r, h = 0.0, site
    while h < alt:
        n = 1.0 + _n0(lam) * math.exp(-h / 8000.0)
        r += sz / math.sqrt(n * n - sz * sz) - sz / math.sqrt(1.0 - sz * sz)
        h += 1.0
    return r


CONF = """TURBULENCE_REF_WAVEL 0.5
TURBULENCE_SEEING 0.8
TURBULENCE_PROF_FILE prof.txt
ZENITH_ANGLE {zangle}
SOURCE_XPOS 0.00001
SOURCE_YPOS -0.00002
WFOUTPUT 1
WF_FILE_PREFIX wf_
SHM_OUTPUT 0
MAKE_SWAVEFRONT 1
SWF_WRITE2DISK 1
SWF_FILE_PREFIX swf_
SHM_SOUTPUT 1
SHM_SPREFIX {prefix}
SHM_SOUTPUTM 0
PUPIL_SCALE 0.02
WFsize 32
WFTIME_STEP 0.001
TIME_SPAN 0.003
NB_TSPAN 1
WF_RAW_SIZE 64
MASTER_SIZE 256
WAVEFRONT_AMPLITUDE 1
FRESNEL_PROPAGATION 1
FRESNEL_PROPAGATION_BIN 1000.0
ATSIM_NOISE_SEED 4321
"""
PROFILE = """# alt cn2 speed dir L0 l0 sigmaW Lwind
0      0.5  10.0  0.0   20.0 0.01 0 0
5000   0.3  20.0  1.57  25.0 0.01 0 0
10000  0.2  30.0  3.49  30.0 0.01 0 0
"""


def _run_cli(cwd, args, env):
    p = subprocess.run([CLI] + args, cwd=cwd, env=env, capture_output=True, text=True, timeout=120)
    sys.stdout.write(p.stdout[-2000:])
    sys.stderr.write(p.stderr[-2000:])
    return p


@pytest.mark.parametrize("zangle", [0.0, 0.3])
def test_mkwfs_end_to_end_vs_oracle(oracle, tmp_path, zangle):
    prefix = f"atsimt{os.getpid()}z{int(zangle * 10)}_"
    (tmp_path / "WFsim.conf").write_text(CONF.format(zangle=zangle, prefix=prefix))
    (tmp_path / "prof.txt").write_text(PROFILE)
    env = dict(os.environ, ATSIM_SHM_DIR=str(tmp_path))
    print(f"mkwfs zangle={zangle} N=64 Nout=32 M=256 L=3 W=2", flush=True)
    try:
        p = _run_cli(tmp_path, ["mkwfs", "1.65", "0"], env)
        assert p.returncode == 0
        M, N, Nout, nf = 256, 64, 32, 3
        dx, dt, lref = f32(0.02), f32(0.001), f32(0.5 * 0.000001)
        slam = 1.0e-6 * f32(1.65)
        z = f32(zangle)
        alts = np.array([0.0, 5000.0, 10000.0])
        cn2 = np.array([0.5, 0.3, 0.2])
        cn2 = cn2 * (1.0 / cn2.sum())
        speeds, dirs = [10.0, 20.0, 30.0], [0.0, 1.57, 3.49]
        # the reference normalisation of the masters it wrote
        masters = [fitsutil.read(str(tmp_path / f"t{l:03d}_{M}_f.fits"))[0].astype(np.float32) for l in range(3)]
        r0 = np.mean([oracle.r0_measure(m.astype(np.float64), 50) for m in masters])
        r0goal = lref / (f32(0.8) / 3600.0 / 180.0 * math.pi) / dx
        tmp1 = (r0 / r0goal) ** (5.0 / 6.0)
        for l in range(3):
            m = (masters[l].astype(np.float64) * tmp1).astype(np.float32)
            masters[l] = (m.astype(np.float64) * math.sqrt(cn2[l] / math.cos(z))).astype(np.float32)
        sa, sli = oracle.superlayers(alts, cn2, f32(1000.0))
        sx, sy = f32(0.00001), f32(-0.00002)
        x0 = [0.05 * M + sx * a / dx for a in alts]
        y0 = [0.05 * M + (_roffset(0.0, a, z, slam) - _roffset(0.0, a, z, lref)) / dx + sy * a / dx for a in alts]
        vx = [s * math.cos(d) / dx for s, d in zip(speeds, dirs)]
        vy = [s * math.sin(d) / dx for s, d in zip(speeds, dirs)]
        sc = lref / slam * _n0(slam) / _n0(lref)
        scoeff = np.array([[1.0] * 3, [sc] * 3])
        os_ = oracle.Sim(N, Nout, np.stack(masters).astype(np.float64), sa, sli, [lref, slam], scoeff, [0, 1],
                         x0, y0, vx, vy, dx, dt, z, True)
        lpm = int(1.0e12 * slam + 0.5)
        cubes = {}
        for key, pre in (("ref", "wf_"), ("sci", "swf_")):
            for q in ("pha", "amp"):
                a, cards = fitsutil.read(str(tmp_path / f"{pre}{0:08d}.{lpm:09d}.{q}.fits"))
                assert int(cards["BITPIX"]) == -32 and a.shape == (nf, Nout, Nout)
                cubes[key, q] = a
        worst = {"amp": 0.0, "pha": 0.0, "mism": 0}
        for k in range(nf):
            _, _, pha_o, amp_o = os_.frame(want_field=False)
            for w, key in enumerate(("ref", "sci")):
                worst["amp"] = max(worst["amp"], rel_l2(cubes[key, "amp"][k], amp_o[w]))
                diff = cubes[key, "pha"][k] - pha_o[w]
                bad = np.abs(diff) > 1.0
                worst["mism"] = max(worst["mism"], int(bad.sum()))
                worst["pha"] = max(worst["pha"], rel_l2(cubes[key, "pha"][k][~bad], pha_o[w][~bad]))
        print("worst", worst)
        assert worst["amp"] < 1e-5 and worst["pha"] < 1e-5 and worst["mism"] <= 1e-3 * Nout * Nout
        # science shared-memory image: last frame, counter, TIME keyword
        raw = open(tmp_path / f"{prefix}pha.im.shm", "rb").read()
        assert struct.unpack_from("<q", raw, 176)[0] == nf and struct.unpack_from("<q", raw, 192)[0] == 1
        shm = np.frombuffer(raw, "<f4", Nout * Nout, 200).reshape(Nout, Nout)
        assert np.array_equal(shm, cubes["sci", "pha"][nf - 1].astype(np.float32))
        kw = 200 + 4 * Nout * Nout
        assert raw[kw:kw + 4] == b"TIME" and raw[kw + 16:kw + 17] == b"D"
        assert abs(struct.unpack_from("<d", raw, kw + 24)[0] - (nf - 1) * dt) < 1e-12
        # a second run finds its masters and, with SKIP_EXISTING, the written span
        with open(tmp_path / "WFsim.conf", "a") as f:
            f.write("SKIP_EXISTING 1\n")
        p = _run_cli(tmp_path, ["mkwfs", "1.65", "0"], env)
        assert p.returncode == 0
        assert "3 turbulence master files" in p.stdout and "Start TSPAN = 1" in p.stdout
    finally:
        for q in ("pha", "amp"):
            for s in range(10):
                sp = f"/dev/shm/sem.{prefix}{q}_sem{s:02d}"
                if os.path.exists(sp):
                    os.unlink(sp)


@pytest.fixture(scope="module")
def clib():
    from atmosphericturbulencesimul_amd._lib import LIB_PATH
    L = C.CDLL(LIB_PATH)
    L.atsim_compat_exit_on_error(0)
    for f in ("create_2DCimage_ID", "create_2Dimage_ID"):
        getattr(L, f).restype = C.c_long
        getattr(L, f).argtypes = [C.c_char_p, C.c_long, C.c_long]
    L.image_ID.restype = C.c_long
    L.image_ID.argtypes = [C.c_char_p]
    L.delete_image_ID.argtypes = [C.c_char_p]
    L.atsim_image_data.restype = C.c_void_p
    L.atsim_image_data.argtypes = [C.c_long]
    L.do2dfft.restype = C.c_long
    L.do2dfft.argtypes = [C.c_char_p, C.c_char_p]
    L.do2dffti.restype = C.c_long
    L.do2dffti.argtypes = [C.c_char_p, C.c_char_p]
    L.permut.argtypes = [C.c_char_p]
    L.Fresnel_propagate_wavefront.argtypes = [C.c_char_p, C.c_char_p, C.c_double, C.c_double, C.c_double]
    L.Fresnel_propagate_cube.restype = C.c_long
    L.Fresnel_propagate_cube.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.c_double, C.c_double,
                                         C.c_long, C.c_double]
    L.make_master_turbulence_screen.argtypes = [C.c_char_p, C.c_char_p, C.c_long, C.c_float, C.c_float, C.c_long]
    L.atsim_compat_set_seed.argtypes = [C.c_ulonglong]
    return L


def _img(L, name, dtype, shape):
    ID = L.image_ID(name.encode())
    assert ID >= 0, name
    ptr = L.atsim_image_data(ID)
    n = int(np.prod(shape))
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))), shape=(n,)).reshape(shape)


def _complex_image(L, name, n, seed):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    L.delete_image_ID(name.encode())
    ID = L.create_2DCimage_ID(name.encode(), n, n)
    assert ID >= 0
    _img(L, name, np.float32, (n, n, 2))[:] = np.stack([a.real, a.imag], -1)
    return a


def _as_c(L, name, n):
    v = _img(L, name, np.float32, (n, n, 2))
    return v[..., 0].astype(np.float64) + 1j * v[..., 1]


def test_compat_fft_and_permut(clib, oracle):
    n = 128
    print("compat dofft/permut n=128", flush=True)
    a = _complex_image(clib, "cfin", n, 1)
    assert clib.do2dfft(b"cfin", b"cfout") >= 0
    assert rel_l2(_as_c(clib, "cfout", n), oracle.fft2d(a.astype(np.complex128), -1)) < 2e-6
    assert clib.do2dffti(b"cfin", b"cfouti") >= 0
    assert rel_l2(_as_c(clib, "cfouti", n), oracle.fft2d(a.astype(np.complex128), +1)) < 2e-6
    # permut moves elements: bit-exact against the oracle's swap loops
    ref = oracle.permut(a.astype(np.complex128))
    assert clib.permut(b"cfin") == 0
    assert np.array_equal(_as_c(clib, "cfin", n), ref)
    clib.delete_image_ID(b"rimg")
    assert clib.create_2Dimage_ID(b"rimg", n, n) >= 0
    r = np.random.default_rng(2).standard_normal((n, n)).astype(np.float32)
    _img(clib, "rimg", np.float32, (n, n))[:] = r
    assert clib.permut(b"rimg") == 0
    assert np.array_equal(_img(clib, "rimg", np.float32, (n, n)), oracle.permut(r.astype(np.float64)).real)


def test_compat_fresnel_vs_oracle(clib, oracle):
    n, dx, z, lam = 256, 0.01, 3000.0, 0.7e-6
    print("compat fresnelpw n=256", flush=True)
    a = _complex_image(clib, "fwin", n, 3)
    assert clib.Fresnel_propagate_wavefront(b"fwin", b"fwout", dx, z, lam) == 0
    assert rel_l2(_as_c(clib, "fwout", n), oracle.fresnel(a.astype(np.complex128), dx, z, lam)) < 1e-5
    assert clib.Fresnel_propagate_cube(b"fwin", b"fwamp", b"fwpha", dx, 0.0, 5000.0, 4, lam) >= 0
    amp_o, pha_o = oracle.fresnel_cube(a.astype(np.complex128), dx, 0.0, 5000.0, 4, lam)
    amp_g = _img(clib, "fwamp", np.float32, (4, n, n))
    assert rel_l2(amp_g, amp_o) < 1e-5


def test_compat_master_screen_vs_oracle(clib, oracle):
    M, L0px, l0px, seed = 256, 1000.0, 1.0, 987654
    print("compat mkmastert M=256", flush=True)
    clib.atsim_compat_set_seed(seed)
    assert clib.make_master_turbulence_screen(b"tm1", b"tm2", M, L0px, l0px, 0) == 0
    _, z32 = oracle.noise_c64(seed, 0, M * M, f32=True)
    o1, o2, _, _ = oracle.make_master(z32.reshape(M, M).astype(np.complex128), f32(L0px), f32(l0px))
    assert rel_l2(_img(clib, "tm1", np.float32, (M, M)), o1) < 1e-5
    assert rel_l2(_img(clib, "tm2", np.float32, (M, M)), o2) < 1e-5
    # the next screen uses the next noise stream
    assert clib.make_master_turbulence_screen(b"tm3", b"tm4", M, L0px, l0px, 0) == 0
    _, z32 = oracle.noise_c64(seed, 1, M * M, f32=True)
    o1, _, _, _ = oracle.make_master(z32.reshape(M, M).astype(np.complex128), f32(L0px), f32(l0px))
    assert rel_l2(_img(clib, "tm3", np.float32, (M, M)), o1) < 1e-5
