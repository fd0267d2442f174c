"""Build libxpir_amd.so (HIP kernels + C ABI) in-tree for gfx950.

    python -m xpir_amd.build            # compile every HIP source, link the library

The library lands next to this file (xpir_amd/libxpir_amd.so) so that it travels with
the source tree; nothing is installed anywhere else.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libxpir_amd.so")
ARCH = os.environ.get("XPIR_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["capi.hip", "import.hip", "mac.hip", "mx.hip", "server.hip", "decrypt.hip", "tune.hip"]

# -packed-fp32-ops: keep the compiler from forming packed fp32 VALU ops (none of this
# code is fp32, the flag is a standing rule for every HIP file built here).
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include")]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    # the C++ facade (reference-shaped server API and session protocol), the server
    # program, and the end-to-end demo program
    facade = os.path.join(HERE, "libpir_amd.so")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", facade + ".tmp",
           os.path.join(HERE, "cpp", "libpir_amd.cpp"), os.path.join(HERE, "cpp", "pir_session.cpp"),
           "-L", HERE, "-lxpir_amd", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    os.replace(facade + ".tmp", facade)
    server = os.path.join(OBJ, "pir_server")
    cmd = ["g++", "-O2", "-std=c++17", "-pthread", "-o", server, os.path.join(HERE, "cpp", "pir_server.cpp"),
           "-L", HERE, "-lpir_amd", "-lxpir_amd", "-Wl,-rpath,$ORIGIN/.."]
    _run(cmd)
    demo = os.path.join(OBJ, "simple_pir")
    cmd = ["g++", "-O2", "-std=c++17", "-o", demo, os.path.join(ROOT, "tests", "cpp", "simple_pir.cpp"),
           "-L", HERE, "-lpir_amd", "-lxpir_amd", "-Wl,-rpath,$ORIGIN/.."]
    _run(cmd)
    if verbose:
        print("built", LIB, facade, server, demo)
    return LIB


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]))


if __name__ == "__main__":
    build(verbose=True)
