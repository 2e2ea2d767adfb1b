"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer:
the library's host C++ (table reader and writer, grids, parameter checks,
the C ABI's error paths) and the oracle, built with -fsanitize by g++
(radchan_amd.build.build_sanitized), run the CPU tests of the ABI, the
tables and the oracle pins in a child process with the sanitizer runtimes
preloaded.  Any invalid access or undefined behaviour aborts that process."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


@pytest.mark.skipif(not (_runtime("libasan.so") and _runtime("libubsan.so")), reason="no gcc sanitizer runtimes")
def test_host_code_under_sanitizers(tmp_path):
    from radchan_amd import build as B
    lib, olib = B.build_sanitized(str(tmp_path))
    for so in (lib, olib):          # instrumented: the sanitizers' hooks are imported
        syms = subprocess.run(["nm", "-D", "--undefined-only", so], capture_output=True, text=True).stdout
        assert "__asan_report" in syms and "__ubsan_handle" in syms, so
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": _runtime("libasan.so") + ":" + _runtime("libubsan.so"),
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:halt_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "RADCHAN_LIB": lib,
        "RADCHAN_ORACLE_LIB": olib,
        "OMP_NUM_THREADS": "4",
    })
    tests = ["tests/test_abi.py", "tests/test_oracle_pins.py", "tests/test_oracle_closed_forms.py",
             "tests/test_estimator_modes.py", "tests/test_zz_no_alloc.py"]
    # (the header shim test links a C program against the in-tree library by name: not run here)
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                          "--deselect", "tests/test_abi.py::test_param_shim_against_reference_headers", *tests],
                         cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
