"""CPU checks of the reference-compatible host layer in libatsim.so: the
configuration reader (read_config_parameter, COREMOD_tools.c:629-690), the
profile reader (AtmosphericTurbulence.c:1945-2004), the image registry, the
FITS writer/reader against an independent numpy FITS codec, the shared-memory
image layout (IMAGE_METADATA, CLIcore.h:179-201) and the command-line argument
checks (CLIcore.c:1630-1853).  None of these calls reaches the GPU.

The configuration and profile fixtures under tests/golden/ are the
reference's own WFsim.conf and turbul.prof data files."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from tests import fitsutil

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def lib():
    from atmosphericturbulencesimul_amd._lib import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.fail(f"{LIB_PATH} not built")
    L = C.CDLL(LIB_PATH)
    L.atsim_compat_exit_on_error(0)
    L.create_image_ID.restype = C.c_long
    L.create_image_ID.argtypes = [C.c_char_p, C.c_long, C.POINTER(C.c_long), C.c_int, C.c_int, C.c_int]
    for f in ("create_2Dimage_ID", "create_2Dimage_ID_double"):
        getattr(L, f).restype = C.c_long
        getattr(L, f).argtypes = [C.c_char_p, C.c_long, C.c_long]
    for f in ("create_3Dimage_ID", "create_3Dimage_ID_double"):
        getattr(L, f).restype = C.c_long
        getattr(L, f).argtypes = [C.c_char_p, C.c_long, C.c_long, C.c_long]
    L.image_ID.restype = C.c_long
    L.image_ID.argtypes = [C.c_char_p]
    L.delete_image_ID.restype = C.c_long
    L.delete_image_ID.argtypes = [C.c_char_p]
    L.chname_image_ID.restype = C.c_long
    L.chname_image_ID.argtypes = [C.c_char_p, C.c_char_p]
    L.atsim_image_data.restype = C.c_void_p
    L.atsim_image_data.argtypes = [C.c_long]
    L.atsim_image_type.argtypes = [C.c_long]
    L.atsim_image_size.restype = C.c_long
    L.atsim_image_size.argtypes = [C.c_long, C.c_int]
    L.load_fits.restype = C.c_long
    L.load_fits.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    L.save_fl_fits.argtypes = [C.c_char_p, C.c_char_p]
    L.save_db_fits.argtypes = [C.c_char_p, C.c_char_p]
    L.read_config_parameter.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
    L.read_config_parameter_exists.argtypes = [C.c_char_p, C.c_char_p]
    L.atsim_cli_execute.argtypes = [C.c_char_p]
    L.atsim_compat_set_shm_dir.argtypes = [C.c_char_p]
    L.file_exists.argtypes = [C.c_char_p]
    return L


def _b(s):
    return s.encode()


def _array(lib, ID, dtype, n):
    ptr = lib.atsim_image_data(ID)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))), shape=(n,))


def _conf(lib, path, key):
    buf = C.create_string_buffer(256)
    rc = lib.read_config_parameter(_b(path), _b(key), buf)
    return rc, buf.value.decode()


def test_read_config_parameter_on_reference_conf(lib):
    f = os.path.join(GOLDEN, "WFsim.conf")
    assert _conf(lib, f, "TURBULENCE_REF_WAVEL") == (1, "0.500000")
    assert _conf(lib, f, "TURBULENCE_PROF_FILE") == (1, "turbul.prof")
    assert _conf(lib, f, "WFsize") == (1, "64")
    assert _conf(lib, f, "FRESNEL_PROPAGATION_BIN") == (1, "1000.0")
    assert _conf(lib, f, "WF_FILE_PREFIX") == (1, "./WF_files/wf_")
    # absent keys read as "-" (atof -> 0) and report 0
    assert _conf(lib, f, "SITE_ALT") == (0, "-")
    assert lib.read_config_parameter_exists(_b(f), b"SKIP_EXISTING") == 1
    assert lib.read_config_parameter_exists(_b(f), b"REALTIME") == 0


def test_read_config_parameter_last_line_wins(lib, tmp_path):
    f = tmp_path / "c.conf"
    f.write_text("KEY 1\nOTHER 5 comment words\nKEY 2 trailing\n# KEY 9\nEMPTY\n")
    assert _conf(lib, str(f), "KEY") == (1, "2")
    assert _conf(lib, str(f), "OTHER") == (1, "5")
    assert _conf(lib, str(f), "EMPTY") == (1, "")


class _Profile(C.Structure):
    _fields_ = [("n", C.c_long)] + [(k, C.POINTER(C.c_double)) for k in
                                     ("alt", "cn2", "speed", "dir", "outer", "inner", "sigmaw", "lwind")]


def _profile(lib, path, site):
    p = _Profile()
    lib.atsim_profile_read.argtypes = [C.c_char_p, C.c_double, C.POINTER(_Profile)]
    assert lib.atsim_profile_read(_b(path), site, C.byref(p)) == 0
    out = {k: np.array([getattr(p, k)[i] for i in range(p.n)]) for k, _ in _Profile._fields_[1:]}
    lib.atsim_profile_free.argtypes = [C.POINTER(_Profile)]
    lib.atsim_profile_free(C.byref(p))
    return out


def test_profile_reader_reference_profile(lib):
    f = os.path.join(GOLDEN, "turbul.prof")
    p = _profile(lib, f, 0.0)
    assert p["alt"].tolist() == [4215, 4230, 4349, 5007, 12000, 16200, 23701]
    assert p["cn2"].tolist() == [5.32, 1.47, 1.08, 2.11, 1.83, 1.48, 0.697]
    assert p["dir"][-1] == 3.3 and p["speed"][4] == 22.0
    # the 4-column file has no scales: they read as 0
    assert not p["outer"].any() and not p["inner"].any()
    # layers below SITE_ALT - 0.1 are dropped (AtmosphericTurbulence.c:1988)
    p = _profile(lib, f, 4349.05)
    assert p["alt"].tolist() == [4349, 5007, 12000, 16200, 23701]


def test_profile_reader_eight_columns(lib, tmp_path):
    f = tmp_path / "p.prof"
    f.write_text("# alt cn2 v dir L0 l0 sw Lw\n\n100 0.5 10 0.1 25 0.01 1 30\n"
                 "  2000 0.5 20 1.5 30 0.02 2 40  extra\n")
    p = _profile(lib, str(f), 0.0)
    assert p["alt"].tolist() == [100, 2000]
    assert p["outer"].tolist() == [25, 30] and p["inner"].tolist() == [0.01, 0.02]
    assert p["lwind"].tolist() == [30, 40]


def test_registry_create_rename_delete(lib):
    ID = lib.create_2Dimage_ID(b"rgA", 5, 3)
    assert ID >= 0 and lib.image_ID(b"rgA") == ID
    assert lib.atsim_image_size(ID, 0) == 5 and lib.atsim_image_size(ID, 1) == 3
    a = _array(lib, ID, np.float32, 15)
    assert not a.any()  # zero-initialised
    a[:] = np.arange(15)
    # creating an existing name returns the same image
    assert lib.create_2Dimage_ID(b"rgA", 5, 3) == ID
    assert lib.chname_image_ID(b"rgA", b"rgB") == ID
    assert lib.image_ID(b"rgA") == -1 and lib.image_ID(b"rgB") == ID
    assert _array(lib, ID, np.float32, 15)[7] == 7
    assert lib.delete_image_ID(b"rgB") == ID and lib.image_ID(b"rgB") == -1
    assert lib.delete_image_ID(b"rgB") == -1


@pytest.mark.parametrize("dbl", [False, True])
def test_fits_round_trip_and_layout(lib, tmp_path, dbl):
    nx, ny, nz = 7, 5, 3
    ID = (lib.create_3Dimage_ID_double if dbl else lib.create_3Dimage_ID)(b"ft", nx, ny, nz)
    dt = np.float64 if dbl else np.float32
    ref = np.random.default_rng(1).standard_normal(nx * ny * nz).astype(dt)
    _array(lib, ID, dt, ref.size)[:] = ref
    path = str(tmp_path / "a.fits")
    save = lib.save_db_fits if dbl else lib.save_fl_fits
    assert save(b"ft", _b(path)) == 0
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    arr, cards = fitsutil.read(path)
    assert int(cards["BITPIX"]) == (-64 if dbl else -32)
    assert arr.shape == (nz, ny, nx)  # NAXIS1 = x, the fastest axis
    assert np.array_equal(arr.ravel(), ref.astype(np.float64))
    # without '!', an existing file is kept; with '!', it is replaced
    _array(lib, ID, dt, ref.size)[:] = 0
    save(b"ft", _b(path))
    assert np.array_equal(fitsutil.read(path)[0].ravel(), ref.astype(np.float64))
    assert save(b"ft", _b("!" + path)) == 0
    assert not fitsutil.read(path)[0].any()
    lib.delete_image_ID(b"ft")
    # reader: back into a registry image of the same type
    fitsutil.write(path, ref.reshape(nz, ny, nx), -64 if dbl else -32)
    ID = lib.load_fits(_b(path), b"ld", 1)
    assert ID >= 0 and lib.atsim_image_type(ID) == (4 if dbl else 3)
    assert np.array_equal(_array(lib, ID, dt, ref.size), ref)
    lib.delete_image_ID(b"ld")


@pytest.mark.parametrize("bitpix", [8, 16, 32])
def test_fits_reader_integer_images(lib, tmp_path, bitpix):
    ref = np.arange(24).reshape(4, 6) * (3 if bitpix > 8 else 1) - (0 if bitpix == 8 else 20)
    path = str(tmp_path / "i.fits")
    fitsutil.write(path, ref, bitpix)
    ID = lib.load_fits(_b(path), b"li", 1)
    assert ID >= 0 and lib.atsim_image_type(ID) == 3
    assert np.array_equal(_array(lib, ID, np.float32, ref.size), ref.ravel().astype(np.float32))
    lib.delete_image_ID(b"li")
    assert lib.load_fits(_b(str(tmp_path / "missing.fits")), b"x", 0) == -1
    assert lib.file_exists(_b(path)) == 1 and lib.file_exists(b"/nonexistent/q.fits") == 0


def test_shared_image_layout(lib, tmp_path):
    name = f"atsimtest{os.getpid()}"
    lib.atsim_compat_set_shm_dir(_b(str(tmp_path)))
    try:
        size = (C.c_long * 2)(6, 4)
        ID = lib.create_image_ID(_b(name), 2, size, 3, 1, 2)
        assert ID >= 0
        a = _array(lib, ID, np.float32, 24)
        a[:] = np.arange(24, dtype=np.float32) + 0.5
        raw = open(tmp_path / f"{name}.im.shm", "rb").read()
        # IMAGE_METADATA: 200 bytes, then the data, then NBkw 120-byte keywords
        assert len(raw) == 200 + 24 * 4 + 2 * 120
        assert raw[0:80].split(b"\0")[0] == _b(name)
        naxis, s0, s1, s2, nel = struct.unpack_from("<5q", raw, 80)
        assert (naxis, s0, s1, s2, nel) == (2, 6, 4, 0, 24)
        assert struct.unpack_from("<i", raw, 120)[0] == 3  # atype
        assert struct.unpack_from("<3i", raw, 160) == (1, 0, 0)  # shared, write, status
        assert struct.unpack_from("<3q", raw, 176) == (0, 0, 2)  # cnt0, cnt1, NBkw
        assert np.array_equal(np.frombuffer(raw, "<f4", 24, 200), a)
        assert raw[200 + 96 + 16:200 + 96 + 17] == b"N"  # unused keyword type
        for s in range(10):
            assert os.path.exists(f"/dev/shm/sem.{name}_sem{s:02d}")
    finally:
        lib.delete_image_ID(_b(name))
        for s in range(10):
            p = f"/dev/shm/sem.{name}_sem{s:02d}"
            if os.path.exists(p):
                os.unlink(p)
        lib.atsim_compat_set_shm_dir(b"/tmp")


def test_cli_argument_checks(lib, tmp_path, capfd):
    ex = lib.atsim_cli_execute
    assert ex(b"") == 0 and ex(b"   # comment only") == 0
    assert ex(b"exit") == 1 and ex(b"quit") == 1
    assert ex(b"nosuchcommand 1 2") == -1
    assert ex(b"mkwfs abc 1") == -1       # lambda must be a number
    assert ex(b"mkwfs 1.65") == -1        # missing argument
    assert ex(b"fresnelpw 1.0 out 0.01 1000 5e-7") == -1  # input must be an image name
    lib.create_2Dimage_ID(b"clim", 4, 4)
    assert ex(b"fresnelpw in clim 0.01 1000 5e-7") == -1  # output must not be an image
    assert ex(b"mkmastert a b 2.5x 10 0.1") == -1
    assert ex(b"listim") == 0
    path = str(tmp_path / "c.fits")
    assert ex(_b(f"savefits clim {path}")) == 0 and os.path.exists(path)
    assert ex(_b(f"loadfits {path} clim2")) == 0 and lib.image_ID(b"clim2") >= 0
    assert ex(b"rm clim clim2") == 0 and lib.image_ID(b"clim") == -1
    C.CDLL(None).fflush(None)  # C stdio buffers the messages
    out = capfd.readouterr().out
    assert "should be" in out and "unknown command" in out
