"""Build libdmri.so (gfx950 HIP kernels + C++ host code) and the dmri-solve CLI in-tree.

    python dmri-fem-cloud_amd/build_native.py            # build everything
    python dmri-fem-cloud_amd/build_native.py --check    # print the output paths
    python dmri-fem-cloud_amd/build_native.py --sanitized OUT  # the CLI under ASan + UBSan (host code)

Outputs: dmri-fem-cloud_amd/lib/libdmri.so and dmri-fem-cloud_amd/lib/dmri-solve.
Object files go to dmri-fem-cloud_amd/_build/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("DMRI_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    # packed FP32 instructions stay disabled for every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC,
]

LIB_SOURCES = ["dmri_kernels.hip", "dmri_api.hip"]
CLI_SOURCE = "dmri_solve_cli.cpp"


def libdmri_path() -> str:
    return os.path.join(LIBDIR, "libdmri.so")


def cli_path() -> str:
    return os.path.join(LIBDIR, "dmri-solve")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in LIB_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = libdmri_path() + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, libdmri_path())
    cli_src = os.path.join(CSRC, CLI_SOURCE)
    if os.path.exists(cli_src):
        tmp = cli_path() + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, cli_src, "-o", tmp, "-L", LIBDIR, "-ldmri",
              f"-Wl,-rpath,$ORIGIN", "-lm"])
        os.replace(tmp, cli_path())
    return libdmri_path()


def build_cli_sanitized(out: str) -> str:
    """The dmri-solve CLI built with AddressSanitizer and UndefinedBehaviorSanitizer (host code only:
    the CLI is plain C++ built by g++, linked against the normal libdmri.so), for running its
    argument parsing and mesh/phase readers under the sanitizers; a finding aborts the program."""
    cli_src = os.path.join(CSRC, CLI_SOURCE)
    _run(["g++", "-O1", "-g", "-std=c++17", "-Wall", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
          "-fno-omit-frame-pointer", "-I", INCLUDE, cli_src, "-o", out, "-L", LIBDIR, "-ldmri",
          f"-Wl,-rpath,{LIBDIR}", "-lm"])
    return out


if __name__ == "__main__":
    if "--sanitized" in sys.argv:
        print(build_cli_sanitized(sys.argv[sys.argv.index("--sanitized") + 1]))
    elif "--check" in sys.argv:
        print(libdmri_path(), cli_path())
    else:
        print(build(verbose=True))
