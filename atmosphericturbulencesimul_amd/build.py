"""Build libatsim.so (HIP kernels + C host library) in-tree for gfx950.

    python -m atmosphericturbulencesimul_amd.build            # build
    python -m atmosphericturbulencesimul_amd.build --clean    # rebuild from scratch

Objects go to atmosphericturbulencesimul_amd/_build/, the shared library to
atmosphericturbulencesimul_amd/libatsim.so and the CLI driver to
atmosphericturbulencesimul_amd/atsim (both loaded/run in place).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libatsim.so")
CLI = os.path.join(PKG, "atsim")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ATSIM_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", INCLUDE, "-I", CSRC,
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-format-truncation", "-I", INCLUDE, "-I", HOST]


def _sources():
    hip = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    host = []
    if os.path.isdir(HOST):
        host = sorted(os.path.join(HOST, f) for f in os.listdir(HOST)
                      if f.endswith(".c") and f != "atsim_cli.c")
    return hip, host


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build(clean=False, verbose=False):
    if clean and os.path.isdir(BUILD):
        shutil.rmtree(BUILD)
    os.makedirs(BUILD, exist_ok=True)
    hip, host = _sources()
    jobs = []
    objs = []
    for src in hip:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        objs.append(obj)
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    for src in host:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        objs.append(obj)
        jobs.append(["gcc"] + C_FLAGS + ["-c", src, "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(HOST, "atsim_cli.c")
    if os.path.exists(cli_src):
        _run(["gcc"] + C_FLAGS + ["-o", CLI, cli_src, "-L", PKG, "-latsim",
                                  "-Wl,-rpath,$ORIGIN", "-lm"])
    return LIB


if __name__ == "__main__":
    build(clean="--clean" in sys.argv, verbose="-v" in sys.argv)
    print(LIB)
