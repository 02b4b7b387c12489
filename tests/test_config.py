"""Loop configuration on the host (no GPU): block sizes the runtime accepts
and reports, and the ones it refuses (common/op_lib_core.c:67-92 reads
OP_BLOCK_SIZE= from argv; the kernels are instantiated for 128, 256 and 512
threads per workgroup)."""

import ctypes as C

import pytest

from oplus2_matlab_amd import op2


@pytest.fixture
def host_runtime():
    L = op2.lib()
    argv = (C.c_char_p * 1)(b"t")
    L.op_init_core(1, argv, 0)
    yield L
    op2.op_exit()
    op2.set_loop_config(None, 0, 0, 0)
    C.c_int.in_dll(L, "OP_block_size").value = 256


def test_block_size_default_and_per_loop(host_runtime):
    assert op2.get_loop_config("res_calc") == {"strategy": 0, "part_size": 0, "block_size": 256}
    op2.set_loop_config("res_calc", op2.STRATEGY_OWNER_GATHER, 512, 128)
    assert op2.get_loop_config("res_calc") == {
        "strategy": op2.STRATEGY_OWNER_GATHER, "part_size": 512, "block_size": 128}
    # other loops keep the default
    assert op2.get_loop_config("update")["block_size"] == 256
    op2.set_loop_config(None, 0, 0, 512)  # every loop
    assert op2.get_loop_config("update")["block_size"] == 512


@pytest.mark.parametrize("bad", [64, 100, 384, 1024, -1])
def test_unsupported_block_size_is_refused(host_runtime, bad):
    with pytest.raises(op2.OP2Error, match="not supported"):
        op2.set_loop_config("res_calc", 0, 0, bad)


def test_op_block_size_from_argv(host_runtime):
    L = host_runtime
    argv = (C.c_char_p * 2)(b"t", b"OP_BLOCK_SIZE=512")
    L.op_init_core(2, argv, 0)
    assert C.c_int.in_dll(L, "OP_block_size").value == 512
    assert op2.get_loop_config("res_calc")["block_size"] == 512
    argv = (C.c_char_p * 2)(b"t", b"OP_BLOCK_SIZE=384")
    L.op2_clear_error()
    L.op_init_core(2, argv, 0)
    assert "not supported" in L.op2_last_error().decode()


@pytest.mark.parametrize("part,block", [(600, 256), (100, 64), (256, 384), (1100, 512)])
def test_invalid_stage_tiles_are_refused(host_runtime, part, block):
    """Stage tiles need a supported block size and at most two cells per
    thread; the setting is left unchanged."""
    with pytest.raises(op2.OP2Error, match="invalid tiles"):
        op2.set_stage_config(part, block)


def test_stage_tiles_accept_defaults_and_res_calc(host_runtime):
    for part, block in [(0, 0), (-1, 0), (512, 256), (128, 128), (1024, 512), (300, 0)]:
        op2.set_stage_config(part, block)
    op2.set_stage_config(0)
