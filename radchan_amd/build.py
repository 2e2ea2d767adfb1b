"""Build the native pieces in-tree.

``build_library()`` compiles the HIP library ``radchan_amd/libradchan.so``
for gfx950 with hipcc.  ``build_oracle()`` compiles the CPU checker
(``oracle/liboracle.so``) and, when the RadChan sources are present, the
reference drivers under ``oracle/_ref``.

Numerics flags: ``-ffp-contract=off`` (no implicit fusion: the FP32 path is
reproduced bit for bit by the CPU checker), packed FP32 instructions disabled,
hardware FP64 atomics.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libradchan.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")

SOURCES = ["rc_tables.cpp", "rc_capi.cpp", "rc_solver.hip", "rc_batched.hip", "rc_ckdist.hip",
           "rc_spline.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return res.stdout


def build_library(verbose: bool = False) -> str:
    """Compile every source for gfx950 and link ``libradchan.so``."""
    hipcc = _hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        out = _run(cmd)
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-L/opt/rocm/lib",
          "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]


def build_sanitized(outdir: str) -> tuple:
    """The library with its host C++ (rc_tables.cpp, rc_capi.cpp) built by
    g++ with AddressSanitizer and UndefinedBehaviorSanitizer, linked with
    the HIP objects of build_library (device code is not instrumented),
    and the oracle built the same way, both into `outdir`: for the CPU
    tests of the host code.  Load them with the sanitizers' runtimes
    preloaded and RADCHAN_LIB / RADCHAN_ORACLE_LIB pointing at them.
    Returns (library, oracle library)."""
    os.makedirs(outdir, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    objs = []
    for src in SOURCES:
        if src.endswith(".hip"):
            obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
            if not os.path.exists(obj):
                build_library()
        else:
            obj = os.path.join(outdir, os.path.splitext(src)[0] + ".o")
            _run(["g++", "-std=c++17", "-fPIC", "-O1", "-g", *SAN_FLAGS, "-D__HIP_PLATFORM_AMD__",
                  "-I/opt/rocm/include", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    lib = os.path.join(outdir, "libradchan_san.so")
    _run(["g++", "-shared", *SAN_FLAGS, "-o", lib, *objs, "-L/opt/rocm/lib", "-lamdhip64",
          "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"])
    olib = os.path.join(outdir, "liboracle_san.so")
    _run(["make", "-C", ORACLE_DIR, "-s", "san", f"SAN_OUT={olib}"])
    return lib, olib


def build_dropin_driver() -> str:
    """The C caller of mc_gpu_ / get_results_ used by the drop-in GPU test
    (tests/dropin/dropin_driver.c), linked against the built library with
    a relative run path so that it loads the in-tree libradchan.so."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "dropin", "dropin_driver.c")
    exe = os.path.join(root, "tests", "dropin", "dropin_driver")
    _run(["gcc", "-O2", "-std=c11", "-Wall", "-I", os.path.join(root, "include"), src, "-L", HERE, "-lradchan",
          "-Wl,-rpath,$ORIGIN/../../radchan_amd", "-o", exe])
    return exe


def build_oracle(with_reference: bool = True) -> None:
    """Compile the CPU checker and, if the RadChan tree is present, its
    reference drivers (outputs only under oracle/_ref)."""
    _run(["make", "-C", ORACLE_DIR, "-s", "liboracle.so"])
    ref = os.environ.get("RADCHAN_REF", "/root/reference")
    if with_reference and os.path.isdir(os.path.join(ref, "src", "montecarlo")):
        _run(["make", "-C", ORACLE_DIR, "-s", "ref", f"REF={ref}"])


if __name__ == "__main__":
    build_oracle()
    print(build_library(verbose=True))
