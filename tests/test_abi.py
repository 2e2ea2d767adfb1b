"""The C ABI without a GPU: every symbol of include/*.h is exported, the
structure layouts seen by C and by the bindings agree, and the documented
error paths return defined errors instead of crashing."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import radchan_amd as R
from radchan_amd import lib as L
from radchan_amd import params as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
REF_PARAM = "/root/reference/src/param"


def declared_functions():
    names = set()
    for h in ("radchan.h",):
        text = open(os.path.join(INC, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return names


def test_every_declared_symbol_is_exported():
    lib = L.load()
    names = declared_functions()
    assert {"mc_gpu_", "get_results_", "rc_create", "rc_run", "rc_trace"} <= names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert names <= exported


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text('''#include <stdio.h>
#include <stddef.h>
#include "radchan.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(rc_params), offsetof(rc_params, seed),
         offsetof(rc_params, maxs), offsetof(rc_params, toll), sizeof(rc_tables), sizeof(rc_ray), sizeof(rc_stats));
  return 0;
}''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", INC, str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert vals == [C.sizeof(L.Params), L.Params.seed.offset, L.Params.maxs.offset, L.Params.toll.offset,
                    C.sizeof(L.TablesC), C.sizeof(L.Ray), C.sizeof(L.Stats)]


def test_reference_defaults():
    p = P.reference_params()
    assert (p.imax, p.jmax, p.kmax, p.p_col, p.photmax, p.nVar) == (406, 576, 2304, 48, 6000, 1)
    assert p.grid_num == 9 and list(p.maxs) == [5] * 8 + [100000]
    assert list(p.maxi) == [406] * 5 + [14] * 3 + [2]
    assert list(p.bd) == [2, 2, 1, 1, 4, 4] and (p.nB, p.nT, p.nQ) == (999, 135, 16)
    assert abs(p.fact_mesh - 1.6) < 1e-15 and p.srt == 1 and p.calcwalls == 0


@pytest.mark.parametrize("change,msg", [
    (lambda p: setattr(p, "kmax", 15), "multiple of p_col"),
    (lambda p: p.bd.__setitem__(0, 4), "x and y faces"),
    (lambda p: p.bd.__setitem__(4, 7), "z faces"),
    (lambda p: p.bd.__setitem__(1, 1), "pairs"),
    (lambda p: p.maxi.__setitem__(1, 7), "divide"),
    (lambda p: setattr(p, "photmax", 0), "photmax"),
    (lambda p: setattr(p, "grid_num", 10), "grid_num"),
    (lambda p: setattr(p, "grid_num", 0), "grid_num"),
    (lambda p: setattr(p, "imax", 0), "must be positive"),
    (lambda p: setattr(p, "nVar", 0), "nVar must be positive"),
    (lambda p: setattr(p, "p_col", 0), "multiple of p_col"),
    (lambda p: p.maxi.__setitem__(0, 8), "level 0 must be the full grid"),
    (lambda p: p.maxj.__setitem__(1, 0), "level sizes must be positive"),
    (lambda p: p.maxs.__setitem__(1, 0), "maxs must be positive"),
    (lambda p: [setattr(p, "jmax", 4096), setattr(p, "kmax", 4096), p.maxj.__setitem__(0, 4096),
                p.maxk.__setitem__(0, 4096)], "grid too large"),
    (lambda p: setattr(p, "photmax", 16385), "photmax above 16384"),
    (lambda p: setattr(p, "Ly", 0.0), "domain lengths"),
    (lambda p: setattr(p, "Lz", float("nan")), "domain lengths"),
    (lambda p: p.bd.__setitem__(2, 0), "x and y faces"),
    (lambda p: setattr(p, "nT", 1), "table sizes"),
    (lambda p: setattr(p, "nB", 4097), "16 bits"),
    (lambda p: setattr(p, "grey", 1), "grey runs use nQ=1"),
    (lambda p: [setattr(p, "grey", 1), setattr(p, "grey_true", 1), setattr(p, "nQ", 1)], "true grey uses nB=1"),
    (lambda p: setattr(p, "toll", 1.0), "toll and tollS"),
    (lambda p: setattr(p, "tollS", 0.0), "toll and tollS"),
    (lambda p: [setattr(p, "calcwalls", 1), setattr(p, "adiabw", 0), setattr(p, "adiabe", 0)],
     "calcwalls needs an adiabatic x wall"),
    (lambda p: setattr(p, "sampler", 2), "sampler"),
    (lambda p: setattr(p, "fp64", 2), "floatpoint"),
])
def test_parameter_validation(change, msg):
    p = P.base_params(16, 16, 16, levels=[(16, 16, 16), (8, 8, 8)], maxs=[5, 100])
    p.p_col = 2
    lib = L.load()
    assert lib.rc_set_params(C.byref(p)) == 0
    change(p)
    rc = lib.rc_set_params(C.byref(p))
    assert rc != 0 and msg in L.last_error()


def test_get_results_without_launch_fails_cleanly():
    R.set_params(P.base_params(8, 8, 8, photmax=8, grey=True))
    res = np.full(1000, 3.0, np.float32)
    var = res.copy()
    R.get_results(res, var)
    assert np.all(res == 3.0) and "without a preceding mc_gpu_" in L.last_error()


def test_schedule_setters_reject_null_contexts():
    """The schedule and refill setters and getters of the batched solver
    return an error with a message for a null context (no GPU needed)."""
    lib = L.load()
    assert lib.rc_set_schedule(None, 20, 4, 5) != 0 and "schedule" in L.last_error()
    assert lib.rc_set_refill(None, 0) != 0 and "rc_set_refill" in L.last_error()
    mode = C.c_int32(7)
    assert lib.rc_get_refill(None, C.byref(mode)) != 0 and mode.value == 7
    assert lib.rc_group_set_schedule(None, 20, 4, 5) != 0
    assert lib.rc_group_set_refill(None, 1) != 0


def test_table_reader_errors(tmp_path):
    with pytest.raises(RuntimeError, match="cannot open"):
        R.Tables.read(str(tmp_path / "none"), 4, 3, 16, 0.5)
    d = tmp_path / "bad"
    d.mkdir()
    (d / "planck-mean.txt").write_text("300 1e-1 0 0\n400 oops\n")
    with pytest.raises(RuntimeError, match="malformed"):
        R.Tables.read(str(d), 4, 3, 16, 0.5)


@pytest.mark.skipif(not os.path.isdir(REF_PARAM), reason="RadChan param headers not present")
def test_param_shim_against_reference_headers(tmp_path):
    """A file that includes the reference's param.h/NarrowBand.h and the shim
    hands the compiled-in configuration to the library at load time."""
    exe = tmp_path / "shim_check"
    subprocess.run(["g++", "-O1", "-w", "-I", REF_PARAM, "-I", INC,
                    os.path.join(ROOT, "tests", "shim", "shim_check.cpp"),
                    os.path.join(ROOT, "tests", "shim", "shim_values.cpp"),
                    "-L", os.path.dirname(L.LIB_PATH), "-lradchan",
                    "-Wl,-rpath," + os.path.dirname(L.LIB_PATH), "-o", str(exe)], check=True)
    v = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    # imax jmax kmax p_col photmax grid_num maxi[5] maxj[8] maxk[8] maxs[8] nB nT nQ bdb srt toll*1e6
    # sampler (random 1: Sobol) fp64 (floatpoint 0: single precision)
    assert v == [406, 576, 2304, 48, 6000, 9, 14, 3, 9, 100000, 999, 135, 16, 4, 1, 1000, 1, 0]
