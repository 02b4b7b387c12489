"""Python binding of the OP2 API over the C ABI of libop2hip.so (ctypes).

Mirrors the reference API (common/op_lib.h, common/op_seq.h) name for name:
op_init, op_decl_set, op_decl_map, op_decl_dat, op_decl_const /
op_decl_const2, op_arg_dat, op_arg_gbl, op_par_loop, op_fetch_data,
op_diagnostic_output, op_timing_output, op_exit, with OP_READ ... OP_MAX and
OP_ID / OP_GBL.  Errors raise OP2Error instead of exiting the process.

Every loop runs on the GPU through the runtime's gfx950 kernels; there is no
host fallback.  The library must have been built (python __graft_entry__.py
build, or make -C oplus2_matlab_amd/csrc).
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libop2hip.so")

OP_READ, OP_WRITE, OP_RW, OP_INC, OP_MIN, OP_MAX = range(6)
OP_ARG_GBL, OP_ARG_DAT = 0, 1
OP_ID = None
OP_GBL = None

PLAN_GENERIC = 16  # plan of the generic loop path (op2_plan_summary_get only)
STRATEGY_AUTO, STRATEGY_COLOURED, STRATEGY_ATOMIC, STRATEGY_DIRECT, STRATEGY_OWNER, STRATEGY_OWNER_GATHER = 0, 1, 2, 3, 4, 5
ERRORS_EXIT, ERRORS_RETURN = 0, 1

_NP_TYPE = {"double": np.float64, "float": np.float32, "int": np.int32,
            "uint": np.uint32, "ll": np.int64, "ull": np.uint64, "bool": np.bool_}


class OP2Error(RuntimeError):
    pass


class _SetCore(C.Structure):
    _fields_ = [("index", C.c_int), ("size", C.c_int), ("name", C.c_char_p)]


class _MapCore(C.Structure):
    _fields_ = [("index", C.c_int), ("from_", C.POINTER(_SetCore)),
                ("to", C.POINTER(_SetCore)), ("dim", C.c_int),
                ("map", C.POINTER(C.c_int)), ("name", C.c_char_p),
                ("map_d", C.c_void_p), ("map_i", C.c_void_p)]


class _DatCore(C.Structure):
    _fields_ = [("index", C.c_int), ("set", C.POINTER(_SetCore)), ("dim", C.c_int),
                ("size", C.c_int), ("data", C.c_void_p), ("data_d", C.c_void_p),
                ("type", C.c_char_p), ("name", C.c_char_p)]


class OpArg(C.Structure):
    """op_arg of common/op_lib_core.h:107-119 (passed by value)."""
    _fields_ = [("index", C.c_int), ("dat", C.POINTER(_DatCore)),
                ("map", C.POINTER(_MapCore)), ("dim", C.c_int), ("idx", C.c_int),
                ("size", C.c_int), ("data", C.c_void_p), ("data_d", C.c_void_p),
                ("type", C.c_char_p), ("acc", C.c_int), ("argtype", C.c_int)]


class PlanSummary(C.Structure):
    _fields_ = [("nblocks", C.c_int), ("ncolors", C.c_int), ("max_block", C.c_int),
                ("max_staged", C.c_int), ("strategy", C.c_int), ("part_size", C.c_int),
                ("mean_thread_colours", C.c_double), ("verify_errors", C.c_longlong),
                ("max_incs", C.c_int), ("cached_plans", C.c_int), ("short_ids", C.c_int)]


class OpPlan(C.Structure):
    """op_plan of op_lib_core.h (common/op_lib_core.h:122-150)."""
    _fields_ = [("name", C.c_char_p), ("set", C.c_void_p), ("nargs", C.c_int),
                ("ninds", C.c_int), ("part_size", C.c_int),
                ("maps", C.POINTER(C.c_void_p)), ("dats", C.POINTER(C.c_void_p)),
                ("idxs", C.POINTER(C.c_int)), ("accs", C.POINTER(C.c_int)),
                ("nthrcol", C.POINTER(C.c_int)), ("thrcol", C.POINTER(C.c_int)),
                ("offset", C.POINTER(C.c_int)), ("ind_maps", C.POINTER(C.POINTER(C.c_int))),
                ("ind_offs", C.POINTER(C.c_int)), ("ind_sizes", C.POINTER(C.c_int)),
                ("nindirect", C.POINTER(C.c_int)), ("loc_maps", C.POINTER(C.POINTER(C.c_short))),
                ("nblocks", C.c_int), ("nelems", C.POINTER(C.c_int)), ("ncolors", C.c_int),
                ("ncolblk", C.POINTER(C.c_int)), ("blkmap", C.POINTER(C.c_int)),
                ("nshared", C.c_int), ("transfer", C.c_float), ("transfer2", C.c_float),
                ("count", C.c_int)]


# every extern "C" entry point of include/op2_hip.h: name -> (restype, argtypes)
_P = C.c_void_p
SIGNATURES = {
    "op_init_core": (None, [C.c_int, C.POINTER(C.c_char_p), C.c_int]),
    "op_decl_set": (_P, [C.c_int, C.c_char_p]),
    "op_decl_map": (_P, [_P, _P, C.c_int, _P, C.c_char_p]),
    "op_decl_dat_core": (_P, [_P, C.c_int, C.c_char_p, C.c_int, _P, C.c_char_p]),
    "op_arg_check": (None, [_P, C.c_int, OpArg, C.POINTER(C.c_int), C.c_char_p]),
    "op_arg_dat": (OpArg, [_P, C.c_int, _P, C.c_int, C.c_char_p, C.c_int]),
    "op_arg_gbl_core": (OpArg, [_P, C.c_int, C.c_char_p, C.c_int]),
    "op_diagnostic_output": (None, []),
    "op_timing_output": (None, []),
    "op_plan_core": (_P, [C.c_char_p, _P, C.c_int, C.c_int, C.POINTER(OpArg), C.c_int,
                          C.POINTER(C.c_int)]),
    "op_timers": (None, [C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "op_timing_realloc": (None, [C.c_int]),
    "op_exit_core": (None, []),
    "op2_init": (None, [C.c_int, C.POINTER(C.c_char_p), C.c_int]),
    "op2_decl_dat_char": (_P, [_P, C.c_int, C.c_char_p, C.c_int, _P, C.c_char_p]),
    "op2_decl_const_char": (None, [C.c_int, C.c_char_p, C.c_int, _P, C.c_char_p]),
    "op2_fetch_data": (None, [_P]),
    "op2_plan_get": (_P, [C.c_char_p, _P, C.c_int, C.c_int, C.POINTER(OpArg), C.c_int,
                          C.POINTER(C.c_int)]),
    "op2_exit": (None, []),
    "op2_par_loop": (C.c_int, [C.c_char_p, _P, C.c_int, C.POINTER(OpArg)]),
    "op2_par_loop_noalloc": (C.c_int, [C.c_char_p, _P, C.c_int, C.POINTER(OpArg)]),
    "op2_set_error_mode": (C.c_int, [C.c_int]),
    "op2_last_error": (C.c_char_p, []),
    "op2_clear_error": (None, []),
    "op2_set_loop_config": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int]),
    "op2_get_loop_config": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "op2_register_loop": (C.c_int, [C.c_char_p, _P]),
    "op2_register_const": (C.c_int, [C.c_char_p, _P]),
    "op2_generic_plan_get": (C.c_int, [C.c_char_p, _P, C.c_int, C.POINTER(OpArg), C.c_int, _P]),
    "op2_generic_owner_plan_get": (C.c_int, [C.c_char_p, _P, C.c_int, C.POINTER(OpArg), C.c_int,
                                             _P, _P]),
    "op2_map_device_ptr": (_P, [_P]),
    "op2_device_scratch": (_P, [C.c_size_t]),
    "op2_report_error": (C.c_int, [C.c_char_p]),
    "op2_checkpoint_record": (C.c_int, [C.c_char_p]),
    "op2_checkpoint_restart": (C.c_int, [C.c_char_p]),
    "op_checkpoint": (C.c_int, []),
    "op2_checkpoint_state": (C.c_int, [C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
    "op2_renumber": (C.c_int, [_P]),
    "op2_set_auto_renumber": (C.c_int, [C.c_int]),
    "op2_renumber_export": (C.c_int, [_P, _P]),
    "op2_autotune_online": (C.c_int, [C.c_char_p, C.c_int]),
    "op2_autotune_online_result": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                             C.POINTER(C.c_double)]),
    "op2_loop_stats": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]),
    "op2_timing_reset": (None, []),
    "op2_stream": (_P, []),
    "op2_sync": (C.c_int, []),
    "op2_dat_device_ptr": (_P, [_P]),
    "op2_plan_summary_get": (C.c_int, [C.c_char_p, _P, C.c_int, C.POINTER(OpArg), C.c_int,
                                       C.c_int, C.POINTER(PlanSummary)]),
    "op2_plan_export": (C.c_int, [C.c_char_p, _P, C.c_int, C.POINTER(OpArg), C.c_int, C.c_int,
                                  _P, _P, _P, _P]),
    "op2_autotune_loop": (C.c_int, [C.c_char_p, _P, C.c_int, C.POINTER(OpArg), C.c_int,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_double)]),
    "op2_set_deferred_reductions": (C.c_int, [C.c_int]),
    "op2_reduction_log": (C.c_int, [_P, C.c_int]),
    "op2_graph_begin": (C.c_int, []),
    "op2_graph_end": (C.c_int, []),
    "op2_graph_launch": (C.c_int, [C.c_int, C.c_int]),
    "op2_graph_destroy": (C.c_int, [C.c_int]),
    "op2_set_graph_fusion": (C.c_int, [C.c_int]),
    "op2_set_stage_config": (C.c_int, [C.c_int, C.c_int]),
    "op2_set_graph_kernel_timing": (C.c_int, [C.c_int]),
    "op2_graph_kernel_times": (C.c_int, [C.c_int, C.c_int, _P, _P]),
    "op2_reduction_log_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
}

_lib = None


def lib():
    """Loads libop2hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OP2Error(f"{LIB_PATH} is missing: build it with "
                           "`python __graft_entry__.py build` (make -C oplus2_matlab_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L.op2_set_error_mode(ERRORS_RETURN)
        _lib = L
    return _lib


def _call(fn, *args):
    L = lib()
    L.op2_clear_error()
    out = getattr(L, fn)(*args)
    msg = L.op2_last_error()
    if msg:
        raise OP2Error(msg.decode(errors="replace"))
    return out


def _b(s):
    return None if s is None else s.encode()


# ---------------------------------------------------------------------------
# handles

class Set:
    def __init__(self, ptr, size, name):
        self._p, self.size, self.name = ptr, size, name
        self._name_b = _b(name)


class Map:
    def __init__(self, ptr, frm, to, dim, data, name):
        self._p, self.from_set, self.to_set, self.dim = ptr, frm, to, dim
        self.data, self.name = data, name


class Dat:
    def __init__(self, ptr, set_, dim, type_, data, name):
        self._p, self.set, self.dim, self.type = ptr, set_, dim, type_
        self.data, self.name = data, name

    def device_ptr(self) -> int:
        return lib().op2_dat_device_ptr(self._p)


_keep = []  # names and arrays the library points at, released by op_exit


def op_init(argv=None, diags: int = 0) -> None:
    """op_init (common/op_lib.cu:128-143)."""
    argv = [b"op2"] + [_b(a) for a in (argv or [])]
    arr = (C.c_char_p * len(argv))(*argv)
    _keep.append(arr)
    _call("op2_init", len(argv), arr, diags)


def op_decl_set(size: int, name: str) -> Set:
    nb = _b(name)
    _keep.append(nb)
    return Set(_call("op_decl_set", size, nb), size, name)


def op_decl_map(frm: Set, to: Set, dim: int, imap, name: str) -> Map:
    a = np.ascontiguousarray(imap, dtype=np.int32).reshape(-1)
    if a.size != dim * frm.size:
        raise OP2Error(f"op_decl_map: {name} has {a.size} entries, expected {dim * frm.size}")
    nb = _b(name)
    _keep.extend([nb, a])
    p = _call("op_decl_map", frm._p, to._p, dim, a.ctypes.data, nb)
    return Map(p, frm, to, dim, a, name)


def op_decl_dat(set_: Set, dim: int, type_: str, data, name: str) -> Dat:
    """op_decl_dat<T> (common/op_lib.h:160-167): checks the element type
    against `type_`, then copies the data to the device."""
    want = _NP_TYPE.get(type_)
    arr = np.asarray(data)
    if want is None or arr.dtype != want:
        raise OP2Error(f'incorrect type specified for dataset "{name}"')
    if not arr.flags.c_contiguous:
        raise OP2Error(f"op_decl_dat: {name} must be C-contiguous")
    if arr.size != dim * set_.size:
        raise OP2Error(f"op_decl_dat: {name} has {arr.size} values, expected {dim * set_.size}")
    nb, tb = _b(name), _b(type_)
    _keep.extend([nb, tb, arr])
    p = _call("op2_decl_dat_char", set_._p, dim, tb, arr.dtype.itemsize, arr.ctypes.data, nb)
    return Dat(p, set_, dim, type_, arr, name)


def op_decl_dat_host(set_: Set, dim: int, type_: str, data, name: str) -> Dat:
    """op_decl_dat_core (common/op_lib_core.c:171-207): registers the dat
    without a device copy (enough for building and checking plans)."""
    arr = np.ascontiguousarray(data, dtype=_NP_TYPE[type_])
    nb, tb = _b(name), _b(type_)
    _keep.extend([nb, tb, arr])
    p = _call("op_decl_dat_core", set_._p, dim, tb, arr.dtype.itemsize, arr.ctypes.data, nb)
    return Dat(p, set_, dim, type_, arr, name)


def op_decl_const2(name: str, dim: int, type_: str, data) -> None:
    want = _NP_TYPE.get(type_)
    arr = np.ascontiguousarray(np.asarray(data, dtype=want)).reshape(-1)
    if arr.size != dim:
        raise OP2Error(f'constant "{name}" has {arr.size} values, expected {dim}')
    _call("op2_decl_const_char", dim, _b(type_), arr.dtype.itemsize, arr.ctypes.data, _b(name))


op_decl_const = op_decl_const2


def op_arg_dat(dat: Dat, idx: int, map_: Map | None, dim: int, type_: str, acc: int) -> OpArg:
    tb = _b(type_)
    _keep.append(tb)
    return lib().op_arg_dat(dat._p, idx, map_._p if map_ is not None else None, dim, tb, acc)


def op_arg_gbl(data: np.ndarray, dim: int, type_: str, acc: int) -> OpArg:
    """op_arg_gbl<T> (common/op_lib.h:184-190); `data` is a numpy array the
    runtime reads (OP_READ) or updates in place (OP_INC / OP_MIN / OP_MAX)."""
    if not isinstance(data, np.ndarray) or not data.flags.c_contiguous:
        raise OP2Error("op_arg_gbl needs a C-contiguous numpy array")
    want = _NP_TYPE.get(type_)
    t = type_ if (want is not None and data.dtype == want) else "error"
    tb = _b(t)
    _keep.extend([tb, data])
    return lib().op_arg_gbl_core(data.ctypes.data, dim, tb, acc)


def op_par_loop(kernel, name: str, set_: Set, *args: OpArg) -> None:
    """op_par_loop (common/op_seq.h): runs loop `name` on the device.  The
    `kernel` argument names the user function, as in the reference; the
    runtime selects its gfx950 kernel by loop name and argument signature."""
    del kernel
    arr = (OpArg * len(args))(*args)
    rc = _call("op2_par_loop", _b(name), set_._p, len(args), arr)
    if rc != 0:
        raise OP2Error(lib().op2_last_error().decode())


def par_loop_noalloc(name: str, set_: Set, *args: OpArg) -> None:
    """op2_par_loop_noalloc: the loop without any device allocation (only
    save_soln, adt_calc and the airfoil / jac1 update loops; see
    include/op2_hip.h)."""
    arr = (OpArg * len(args))(*args)
    rc = _call("op2_par_loop_noalloc", _b(name), set_._p, len(args), arr)
    if rc != 0:
        raise OP2Error(lib().op2_last_error().decode())


def op_fetch_data(dat: Dat) -> np.ndarray:
    """op_fetch_data (common/op_lib.cu:116-121): copies the device data into
    the array given to op_decl_dat and returns it."""
    _call("op2_fetch_data", dat._p)
    return dat.data


def op_diagnostic_output() -> None:
    lib().op_diagnostic_output()


def op_timing_output() -> None:
    _call("op_timing_output")


def op_exit() -> None:
    _call("op2_exit")
    _keep.clear()


# ---------------------------------------------------------------------------
# runtime controls

def set_loop_config(name: str | None, strategy: int = STRATEGY_AUTO, part_size: int = 0,
                    block_size: int = 0) -> None:
    rc = _call("op2_set_loop_config", _b(name), strategy, part_size, block_size)
    if rc != 0:
        raise OP2Error(lib().op2_last_error().decode())


def get_loop_config(name: str) -> dict:
    """The configuration loop `name` runs with: strategy (0 = default), part
    size (0 = default) and block size (threads per workgroup)."""
    st, part, bs = C.c_int(), C.c_int(), C.c_int()
    _call("op2_get_loop_config", _b(name), C.byref(st), C.byref(part), C.byref(bs))
    return {"strategy": st.value, "part_size": part.value, "block_size": bs.value}


def loop_stats(name: str):
    """(count, device seconds, algorithmic bytes per call) or None."""
    c, s, b = C.c_int(), C.c_double(), C.c_double()
    if _call("op2_loop_stats", _b(name), C.byref(c), C.byref(s), C.byref(b)) != 0:
        return None
    return c.value, s.value, b.value


def timing_reset() -> None:
    _call("op2_timing_reset")


def sync() -> None:
    _call("op2_sync")


def stream() -> int:
    return _call("op2_stream")


def autotune(name: str, set_: Set, args, reps: int = 5) -> dict:
    """Times the loop under each candidate strategy and part size on the
    live data (restored afterwards) and keeps the fastest; returns it."""
    arr = (OpArg * len(args))(*args)
    st, part, us = C.c_int(), C.c_int(), C.c_double()
    if _call("op2_autotune_loop", _b(name), set_._p, len(args), arr, reps, C.byref(st),
             C.byref(part), C.byref(us)) != 0:
        raise OP2Error(lib().op2_last_error().decode())
    return {"strategy": st.value, "part_size": part.value,
            "block_size": get_loop_config(name)["block_size"], "us": us.value}


def checkpoint_record(path: str) -> None:
    """Start recording for op_checkpoint (loop counts and reduction values)."""
    if _call("op2_checkpoint_record", _b(path)) != 0:
        raise OP2Error(lib().op2_last_error().decode())


def checkpoint_restart(path: str) -> None:
    """Fast-forward this run to the checkpoint stored in `path`."""
    if _call("op2_checkpoint_restart", _b(path)) != 0:
        raise OP2Error(lib().op2_last_error().decode())


def op_checkpoint() -> bool:
    """Writes a checkpoint; in a restart, returns True at the checkpoint
    where the state was restored."""
    rc = _call("op_checkpoint")
    if rc < 0:
        raise OP2Error(lib().op2_last_error().decode())
    return rc == 1


def checkpoint_state() -> tuple:
    """(mode, loop calls, checkpoints): mode 0 off, 1 recording, 2 forwarding."""
    loops, calls = C.c_ulonglong(), C.c_ulonglong()
    mode = _call("op2_checkpoint_state", C.byref(loops), C.byref(calls))
    return mode, loops.value, calls.value


def renumber(base: "Map") -> None:
    """Renumbers the sets for locality from the target set of `base`
    (internal: fetched data stays in the user's numbering)."""
    if _call("op2_renumber", base._p) != 0:
        raise OP2Error(lib().op2_last_error().decode())


def renumber_export(set_: "Set"):
    """Internal number of every element of the set, or None if the set
    keeps the user's numbering."""
    out = np.zeros(set_.size, np.int32)
    rc = _call("op2_renumber_export", set_._p, out.ctypes.data)
    return out if rc == 1 else None


def autotune_online(name: str, reps: int = 3) -> None:
    """The next calls of loop `name` try each candidate configuration in
    turn (every call a real execution); the fastest is kept."""
    if _call("op2_autotune_online", _b(name), reps) != 0:
        raise OP2Error(lib().op2_last_error().decode())


def autotune_online_result(name: str):
    """The chosen configuration once the online tuning of `name` has
    finished, else None."""
    st, part, us = C.c_int(), C.c_int(), C.c_double()
    rc = _call("op2_autotune_online_result", _b(name), C.byref(st), C.byref(part), C.byref(us))
    if rc < 0:
        raise OP2Error(f"no online tuning of loop {name}")
    if rc != 1:
        return None
    return {"strategy": st.value, "part_size": part.value,
            "block_size": get_loop_config(name)["block_size"], "us": us.value}


def set_deferred_reductions(on: bool) -> bool:
    """Reductions go to the device log instead of the caller's array."""
    return bool(_call("op2_set_deferred_reductions", int(on)))


REDUCTION_LOG_CAPACITY = 1 << 16  # results the device log holds between reads


def reduction_log(max_entries: int = REDUCTION_LOG_CAPACITY) -> np.ndarray:
    """Pending deferred-reduction results, oldest first (waits for the GPU)."""
    out = np.zeros(max_entries, np.float64)
    n = _call("op2_reduction_log", out.ctypes.data, max_entries)
    if n < 0:
        raise OP2Error(lib().op2_last_error().decode())
    return out[:n].copy()


def graph_begin() -> None:
    if _call("op2_graph_begin") != 0:
        raise OP2Error(lib().op2_last_error().decode())


def graph_end() -> int:
    g = _call("op2_graph_end")
    if g < 0:
        raise OP2Error(lib().op2_last_error().decode())
    return g


def graph_launch(gid: int, count: int = 1) -> None:
    if _call("op2_graph_launch", gid, count) != 0:
        raise OP2Error(lib().op2_last_error().decode())


def graph_destroy(gid: int) -> None:
    _call("op2_graph_destroy", gid)


FUSION_OFF, FUSION_LOOPS, FUSION_STAGES = 0, 1, 2


def set_graph_fusion(level) -> int:
    """Loop fusion during graph capture.  FUSION_LOOPS: bres_calc, update,
    the next save_soln and adt_calc run as one kernel, the rms fold rides in
    the next res_calc, and a res that a captured update zeroed is taken as
    zero.  FUSION_STAGES (the default; True means it): also one kernel per
    Runge-Kutta stage, res_calc included, on owner-gather plans that own
    every cell.  q, adt and res keep the separate loops' bits; with stages the
    rms sums are grouped per tile (the same to double rounding).  Returns the
    previous level."""
    lvl = FUSION_STAGES if level is True else FUSION_OFF if level is False else int(level)
    prev = _call("op2_set_graph_fusion", lvl)
    if prev < 0:
        raise OP2Error("op2_set_graph_fusion: not during graph capture")
    return prev


def set_stage_config(part_size: int = 0, block_size: int = 0) -> None:
    """Tiles of the stage kernels (FUSION_STAGES): cells per tile and
    threads (128, 256 or 512, part_size at most twice that).  0: 256 cells
    on 256 threads (the default); -1: res_calc's own configuration.  The
    stage plan is built at graph_begin; res_calc's own launches keep their
    configuration."""
    if _call("op2_set_stage_config", int(part_size), int(block_size)) != 0:
        raise OP2Error(f"op2_set_stage_config: invalid tiles {part_size}/{block_size} "
                       "or graph capture active")


# kinds of graph_kernel_times: the loops a stage kernel runs besides
# res_calc and update, its res traffic, and the form of its adt_calc
KIND_BRES, KIND_SAVE, KIND_ADT, KIND_READS_RES, KIND_STORES_RES = 1, 2, 4, 8, 16
KIND_ADT_SIDES, KIND_SAVE_IN_PLACE, KIND_QOLD_IS_Q = 32, 64, 128


def set_graph_kernel_timing(every: int) -> int:
    """With every = n > 0, the stage kernels numbered n, 2n, ... of each
    capture that follows are bracketed by event-record nodes
    (graph_kernel_times; each node costs a few microseconds of the replay);
    0 turns it off.  Returns the previous setting."""
    prev = _call("op2_set_graph_kernel_timing", max(int(every), 0))
    if prev < 0:
        raise OP2Error("op2_set_graph_kernel_timing: not during graph capture")
    return prev


def graph_kernel_times(gid: int) -> list:
    """[(kind, milliseconds)] of the bracketed kernels of graph gid in its
    last launch, in launch order (waits for the stream)."""
    n = _call("op2_graph_kernel_times", gid, 0, None, None)
    if n < 0:
        raise OP2Error(lib().op2_last_error().decode())
    kinds = (C.c_int * max(n, 1))()
    ms = (C.c_float * max(n, 1))()
    if _call("op2_graph_kernel_times", gid, n, kinds, ms) < 0:
        raise OP2Error(lib().op2_last_error().decode())
    return [(kinds[i], ms[i]) for i in range(n)]


def plan_summary(name: str, set_: Set, args, strategy: int = STRATEGY_COLOURED,
                 part_size: int = 0) -> dict:
    arr = (OpArg * len(args))(*args)
    out = PlanSummary()
    if _call("op2_plan_summary_get", _b(name), set_._p, len(args), arr, strategy, part_size,
             C.byref(out)) != 0:
        raise OP2Error(lib().op2_last_error().decode())
    return {k: getattr(out, k) for k, _ in PlanSummary._fields_}


def plan_export(name: str, set_: Set, args, strategy: int = STRATEGY_COLOURED,
                part_size: int = 0) -> dict:
    """The plan's element order, thread colours, block ranges and block
    colours (built and verified on the host; no GPU needed)."""
    arr = (OpArg * len(args))(*args)
    nb = _call("op2_plan_export", _b(name), set_._p, len(args), arr, strategy, part_size,
               None, None, None, None)
    if nb < 0:
        raise OP2Error(lib().op2_last_error().decode())
    n = set_.size
    perm = np.zeros(n, np.int32)
    thr = np.zeros(n, np.int32)
    off = np.zeros(nb + 1, np.int32)
    col = np.zeros(nb, np.int32)
    _call("op2_plan_export", _b(name), set_._p, len(args), arr, strategy, part_size,
          perm.ctypes.data, thr.ctypes.data, off.ctypes.data, col.ctypes.data)
    return {"perm": perm, "thrcol": thr, "blk_off": off, "blkcol": col}


def op_plan_get(name: str, set_: Set, args, inds, part_size: int = 0) -> dict:
    """op_plan_core's plan in the reference's layout (common/op_lib.cu:155-184
    and op_lib_core.c:627-1089), copied out of the op_plan view: header
    fields, per-block and per-element arrays, ind_maps per indirect dataset
    and loc_maps per argument (None for direct and global arguments)."""
    arr = (OpArg * len(args))(*args)
    ninds = max(inds) + 1 if len(inds) else 0
    iv = (C.c_int * len(inds))(*inds)
    ptr = _call("op_plan_core", _b(name), set_._p, part_size, len(args), arr, ninds, iv)
    if not ptr:
        raise OP2Error(lib().op2_last_error().decode())
    v = OpPlan.from_address(ptr)
    nb, nc, n = v.nblocks, v.ncolors, set_.size

    def arr_of(p, count, dt=np.int32):
        return np.ctypeslib.as_array(p, shape=(count,)).astype(dt).copy() if count else \
            np.zeros(0, dt)

    nind = arr_of(v.nindirect, ninds)
    return {"nblocks": nb, "ncolors": nc, "nshared": v.nshared, "nargs": v.nargs,
            "ninds": v.ninds, "part_size": v.part_size, "count": v.count,
            "transfer": np.float32(v.transfer), "transfer2": np.float32(v.transfer2),
            "nthrcol": arr_of(v.nthrcol, nb), "thrcol": arr_of(v.thrcol, n),
            "offset": arr_of(v.offset, nb), "nelems": arr_of(v.nelems, nb),
            "ncolblk": arr_of(v.ncolblk, nc), "blkmap": arr_of(v.blkmap, nb),
            "ind_offs": arr_of(v.ind_offs, nb * ninds), "ind_sizes": arr_of(v.ind_sizes, nb * ninds),
            "nindirect": nind,
            "ind_maps": [arr_of(v.ind_maps[m], int(nind[m])) for m in range(ninds)],
            "loc_maps": [arr_of(v.loc_maps[a], n, np.int16) if v.loc_maps[a] else None
                         for a in range(len(args))],
            "maps": [v.maps[a] for a in range(len(args))],
            "dats": [v.dats[a] for a in range(len(args))],
            "idxs": arr_of(v.idxs, len(args)), "accs": arr_of(v.accs, len(args))}


# user-kernel names for op_par_loop's first argument (the reference passes
# the host function; only its identity matters there)
save_soln, adt_calc, res_calc, bres_calc, update = (
    "save_soln", "adt_calc", "res_calc", "bres_calc", "update")
res = "res"
