"""The C ABI library: loads, exports every declared entry point, and its host
logic (declaration checks, error mode, plans) works without a GPU."""

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from oplus2_matlab_amd import op2

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HEADER = os.path.join(ROOT, "include", "op2_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*[\w\s\*]+?\b(op2?_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_loads():
    assert op2.lib() is not None


def test_every_declared_function_is_exported():
    names = declared_functions()
    assert "op2_par_loop" in names and "op_decl_set" in names and len(names) >= 30
    L = C.CDLL(op2.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert set(declared_functions()) == set(op2.SIGNATURES)


def test_reference_globals_exported():
    L = C.CDLL(op2.LIB_PATH)
    for g in ("OP_diags", "OP_part_size", "OP_block_size", "OP_set_list", "OP_kernels"):
        C.c_int.in_dll(L, g)


def test_generated_stub_entry_points_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", op2.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for stub in ("op_par_loop_save_soln(char const*, op_set_core*, op_arg, op_arg)",
                 "op_par_loop_res_calc(", "op_par_loop_adt_calc(", "op_par_loop_bres_calc(",
                 "op_par_loop_update(", "op_par_loop_res("):
        assert stub in out, stub


def _init_host():
    L = op2.lib()
    argv = (C.c_char_p * 1)(b"t")
    L.op_init_core(1, argv, 1)


def test_decl_errors_raise():
    _init_host()
    with pytest.raises(op2.OP2Error, match="negative/zero size"):
        op2.op_decl_set(0, "empty")
    a = op2.op_decl_set(4, "a")
    b = op2.op_decl_set(3, "b")
    with pytest.raises(op2.OP2Error, match="invalid data for map"):
        op2.op_decl_map(a, b, 1, np.array([0, 1, 2, 3], np.int32), "bad")
    with pytest.raises(op2.OP2Error, match="incorrect type"):
        op2.op_decl_dat(a, 1, "double", np.zeros(4, np.float32), "wrong")
    m = op2.op_decl_map(a, b, 1, np.array([0, 1, 2, 2], np.int32), "ok")
    d = op2.op_decl_dat_host(b, 2, "double", np.zeros(6), "d")
    # argument checks at OP_diags >= 1 (common/op_lib_core.c:266-305)
    bad = op2.op_arg_dat(d, 1, m, 2, "double", op2.OP_READ)  # idx beyond map dim
    n = C.c_int(0)
    with pytest.raises(op2.OP2Error, match="invalid index"):
        op2._call("op_arg_check", a._p, 0, bad, C.byref(n), b"loop")
    bad = op2.op_arg_dat(d, 0, m, 2, "float", op2.OP_READ)
    with pytest.raises(op2.OP2Error, match="type does not match"):
        op2._call("op_arg_check", a._p, 0, bad, C.byref(n), b"loop")
    op2.lib().op2_exit()


def test_loop_config_validation():
    with pytest.raises(op2.OP2Error):
        op2.set_loop_config("res_calc", 7, 0, 0)
    with pytest.raises(op2.OP2Error):
        op2.set_loop_config("res_calc", 1, 0, 100)
    op2.set_loop_config(None, 0, 0, 0)


def test_online_autotune_arguments():
    op2.lib().op_init_core(0, None, 0)
    with pytest.raises(op2.OP2Error, match="invalid"):
        op2.autotune_online("res_calc", 0)
    with pytest.raises(op2.OP2Error, match="no online tuning"):
        op2.autotune_online_result("never_tuned")
    op2.autotune_online("res_calc", 2)
    assert op2.autotune_online_result("res_calc") is None
    op2.op_exit()


def test_checkpoint_round_trip_host_dats(tmp_path):
    """op_checkpoint writes every dat and a restarted run gets them back at
    the same checkpoint (host-only dats: no GPU needed); a program that
    declares other dats is refused."""
    path = str(tmp_path / "ck.bin")
    L = op2.lib()
    L.op_init_core(0, None, 0)
    cells = op2.op_decl_set(7, "cells")
    a = np.arange(14, dtype=np.float64)
    op2.op_decl_dat_host(cells, 2, "double", a, "a")
    op2.checkpoint_record(path)
    assert op2.op_checkpoint() is False
    op2.op_exit()
    L.op_init_core(0, None, 0)
    cells = op2.op_decl_set(7, "cells")
    b = np.zeros(14)
    op2.op_decl_dat_host(cells, 2, "double", b, "a")
    op2.checkpoint_restart(path)
    assert op2.checkpoint_state()[0] == 2
    assert op2.op_checkpoint() is True
    assert np.array_equal(b, a) and op2.checkpoint_state()[0] == 1
    op2.op_exit()
    L.op_init_core(0, None, 0)
    cells = op2.op_decl_set(7, "cells")
    op2.op_decl_dat_host(cells, 1, "double", np.zeros(7), "a")
    op2.checkpoint_restart(path)
    with pytest.raises(op2.OP2Error, match="does not match"):
        op2.op_checkpoint()
    op2.op_exit()
