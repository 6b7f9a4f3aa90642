"""The C ABI library: it loads, exports every entry point include/stz_amd.h
declares, its structs have the layout the header gives them, and without a
device it fails loudly instead of falling back to the CPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

import stz_amd
from stz_amd import _lib
from oracle import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "stz_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(stz_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared set
    assert sorted(_lib.exported_symbols()) == names


def test_struct_layout_matches_header(tmp_path):
    probe = tmp_path / "probe.c"
    structs = {
        "stz_model_params": _lib.ModelParams,
        "stz_sim_params": _lib.SimParams,
        "stz_mg_config": _lib.MGConfig,
        "stz_solve_info": _lib.SolveInfo,
        "stz_bd_params": _lib.BdParams,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void) {"]
    for cname, py in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines.append("return 0; }")
    probe.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-std=c99", "-o", str(exe), str(probe)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)], text=True).splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got["%s.%s" % (cname, fname)]) == getattr(py, fname).offset, (cname, fname)
    # the oracle's own copy of the parameter structs agrees too
    assert C.sizeof(pyoracle.SimParams) == C.sizeof(_lib.SimParams)
    for fname, _ in _lib.SimParams._fields_:
        assert getattr(pyoracle.SimParams, fname).offset == getattr(_lib.SimParams, fname).offset
    assert C.sizeof(pyoracle.BdParams) == C.sizeof(_lib.BdParams)
    for fname, _ in _lib.BdParams._fields_:
        assert getattr(pyoracle.BdParams, fname).offset == getattr(_lib.BdParams, fname).offset


def test_no_device_fails_loudly():
    if stz_amd.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(stz_amd.StzError) as e:
        stz_amd.ShearSim(stz_amd.problems.simple_shear(16))
    assert e.value.code == _lib.ERR_NO_DEVICE


def test_no_device_fails_loudly_free_boundary():
    if stz_amd.device_count() > 0:
        pytest.skip("a device is visible")
    from stz_amd.bd import BdSim
    with pytest.raises(stz_amd.StzError) as e:
        BdSim(stz_amd.problems.bd_params(16))
    assert e.value.code == _lib.ERR_NO_DEVICE


def test_bad_arguments_rejected_before_device_use():
    lib = _lib.load()
    p = _lib.make_params(stz_amd.problems.simple_shear(2))
    h = C.c_void_p()
    assert lib.stz_sim_create(C.byref(p), 0, C.byref(h)) == _lib.ERR_ARG
    assert b"m >= 4" in lib.stz_last_error()
    p = _lib.make_params(stz_amd.problems.simple_shear(16))
    assert lib.stz_sim_create(C.byref(p), 7, C.byref(h)) == _lib.ERR_ARG
    assert lib.stz_sim_destroy(None) == _lib.OK
    assert lib.stz_sim_set_time(None, 1.0) == _lib.ERR_ARG
    pb = _lib.make_bd_params(stz_amd.problems.bd_params(4))
    assert lib.stz_bd_create(C.byref(pb), C.byref(h)) == _lib.ERR_ARG
    assert b"m >= 8" in lib.stz_last_error()
    assert lib.stz_bd_destroy(None) == _lib.OK
    assert lib.stz_bd_step_quasistatic(None, 1.0, 1) == _lib.ERR_ARG


def test_no_device_fails_loudly_ensemble():
    if stz_amd.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(stz_amd.StzError) as e:
        stz_amd.Ensemble([stz_amd.problems.simple_shear(64)] * 2)
    assert e.value.code == _lib.ERR_NO_DEVICE


def test_ensemble_bad_arguments():
    lib = _lib.load()
    h = C.c_void_p()
    p = _lib.make_params(stz_amd.problems.simple_shear(64))
    assert lib.stz_ensemble_create(C.byref(p), 0, 1, C.byref(h)) == _lib.ERR_ARG
    assert lib.stz_ensemble_create(None, 2, 1, C.byref(h)) == _lib.ERR_ARG
    assert lib.stz_ensemble_destroy(None) == _lib.OK
    assert lib.stz_ensemble_step_quasistatic(None, 1.0, 1) == _lib.ERR_ARG
    n = C.c_int()
    assert lib.stz_ensemble_size(None, C.byref(n)) == _lib.ERR_ARG
