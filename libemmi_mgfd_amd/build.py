"""Build the in-tree shared library libemmi_mgfd.so for gfx950 with hipcc,
then the command-line program: libemmi_cli.so (parameters, survey and data
files, data weighting) and the executable emmi_main, linked against both.

Device code is compiled for gfx950 only; the host-side bit-exact path is
compiled without floating-point contraction.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libemmi_mgfd.so")

HIP_SOURCES = ["mg_kernels.hip", "line_kernels.hip", "solver.hip", "path_kernels.hip", "x0net_kernels.hip"]
HOST_SOURCES = ["capi.cpp", "host_path.cpp", "timing.cpp", "path_api.cpp"]

CLI_LIB_SOURCES = ["params.cpp", "survey.cpp", "weighting.cpp", "inversion.cpp"]
CLI_LIB = os.path.join(HERE, "libemmi_cli.so")
CLI_EXE = os.path.join(HERE, "emmi_main")

ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-lambda-capture"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(args):
    src, obj, flags = args
    cmd = [hipcc()] + flags + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append((os.path.join(CSRC, s), os.path.join(BUILD, s + ".o"), COMMON + DEVICE + ["-x", "hip"]))
    for s in HOST_SOURCES:
        jobs.append((os.path.join(CSRC, s), os.path.join(BUILD, s + ".o"), COMMON + ["-ffp-contract=off"]))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cli()
    if verbose:
        print("built", LIB, CLI_LIB, CLI_EXE)
    return LIB


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_cli() -> None:
    """libemmi_cli.so and emmi_main (host code only, g++)."""
    cli = os.path.join(CSRC, "cli")
    flags = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-pthread", "-I" + os.path.join(ROOT, "include"),
             "-I" + cli]
    objs = []
    for src in CLI_LIB_SOURCES:
        obj = os.path.join(BUILD, "cli_" + src + ".o")
        _run(["g++"] + flags + ["-c", os.path.join(cli, src), "-o", obj])
        objs.append(obj)
    _run(["g++", "-shared", "-o", CLI_LIB + ".tmp"] + objs + ["-lm"])
    os.replace(CLI_LIB + ".tmp", CLI_LIB)
    main_obj = os.path.join(BUILD, "cli_emmi_main.cpp.o")
    _run(["g++"] + flags + ["-c", os.path.join(cli, "emmi_main.cpp"), "-o", main_obj])
    _run(["g++", "-o", CLI_EXE + ".tmp", main_obj, "-L" + HERE, "-lemmi_cli", "-lemmi_mgfd",
          "-Wl,-rpath,$ORIGIN", "-lm", "-pthread"])
    os.replace(CLI_EXE + ".tmp", CLI_EXE)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
