"""The host layer (image registry, shared-memory images, FITS I/O,
configuration and profile readers) built with AddressSanitizer and
UndefinedBehaviorSanitizer by gcc on the CPU and driven by
tests/host_sanitize/driver.c.  Any out-of-bounds access, use after free,
leak or undefined operation found on the way fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "atmosphericturbulencesimul_amd", "host")


def test_host_layer_under_asan_and_ubsan(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = tmp_path / "host_sanitize"
    srcs = [os.path.join(ROOT, "tests", "host_sanitize", "driver.c")] + [
        os.path.join(HOST, f) for f in ("registry.c", "fitsio.c", "conf.c")]
    cmd = [gcc, "-std=gnu11", "-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "include"), "-I", HOST] + srcs + [
           "-o", str(exe), "-lm", "-lpthread", "-lrt"]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-3000:]
    work = tmp_path / "work"
    work.mkdir()
    # verify_asan_link_order=0: the sanitizer runtime need not be the first
    # library of the process when the environment preloads another one
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), str(work), os.path.join(ROOT, "tests", "golden")], capture_output=True, text=True,
                       env=env, timeout=300)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert r.returncode == 0, out[-3000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out and "LeakSanitizer" not in out
    assert "all checks passed" in out
