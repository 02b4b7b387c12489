"""ctypes access to the CPU oracle (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
compare results against, and by bench.py's cpu_baseline leg.  The product library never
imports it.

The oracle restates the reference's sequential order (common/op_seq.h), its
plan builder (common/op_lib_core.c:627-1089) and its OpenMP schedule
(airfoil/<loop>_kernel.cpp, jac1/<loop>_kernel.cpp) in three arithmetic
forms: "f64", "f32" (the reference's own), "mix" (float state and flux,
double update and reductions).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")

MAXARGS = 24
READ, WRITE, RW, INC, MIN, MAX = range(6)

FORMS = {"f64": (np.float64, np.float64), "f32": (np.float32, np.float32),
         "mix": (np.float32, np.float64)}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class PlanIn(C.Structure):
    _fields_ = [("set_size", C.c_int), ("nargs", C.c_int), ("ninds", C.c_int),
                ("part_size", C.c_int),
                ("map", C.c_void_p * MAXARGS), ("map_dim", C.c_int * MAXARGS),
                ("idx", C.c_int * MAXARGS), ("to_size", C.c_int * MAXARGS),
                ("dat_size", C.c_int * MAXARGS), ("dat_dim", C.c_int * MAXARGS),
                ("acc", C.c_int * MAXARGS), ("inds", C.c_int * MAXARGS)]


class Plan(C.Structure):
    _fields_ = [("bsize", C.c_int), ("nblocks", C.c_int), ("ncolors", C.c_int),
                ("nshared", C.c_int), ("ninds", C.c_int), ("nargs", C.c_int),
                ("set_size", C.c_int),
                ("nthrcol", C.POINTER(C.c_int)), ("thrcol", C.POINTER(C.c_int)),
                ("offset", C.POINTER(C.c_int)), ("nelems", C.POINTER(C.c_int)),
                ("ncolblk", C.POINTER(C.c_int)), ("blkmap", C.POINTER(C.c_int)),
                ("blkcol", C.POINTER(C.c_int)),
                ("ind_offs", C.POINTER(C.c_int)), ("ind_sizes", C.POINTER(C.c_int)),
                ("nindirect", C.c_int * MAXARGS),
                ("ind_maps", C.POINTER(C.c_int) * MAXARGS),
                ("loc_maps", C.POINTER(C.c_short) * MAXARGS),
                ("transfer", C.c_float), ("transfer2", C.c_float),
                ("transfer3", C.c_float)]


class PlanArg:
    """One op_arg as the plan builder sees it."""

    def __init__(self, map_=None, map_dim=0, idx=-1, to_size=0, dat_size=0,
                 dat_dim=1, acc=READ, ind=-1):
        self.map = None if map_ is None else np.ascontiguousarray(map_, np.int32)
        self.map_dim, self.idx, self.to_size = map_dim, idx, to_size
        self.dat_size, self.dat_dim, self.acc, self.ind = dat_size, dat_dim, acc, ind


def build_plan(set_size: int, args: list, ninds: int, part_size: int = 0) -> dict:
    """Run the oracle plan builder; returns the plan arrays and check count."""
    pin = PlanIn()
    pin.set_size, pin.nargs, pin.ninds, pin.part_size = set_size, len(args), ninds, part_size
    keep = []
    for i, a in enumerate(args):
        if a.map is not None:
            keep.append(a.map)
            pin.map[i] = a.map.ctypes.data
        pin.map_dim[i], pin.idx[i], pin.to_size[i] = a.map_dim, a.idx, a.to_size
        pin.dat_size[i], pin.dat_dim[i], pin.acc[i], pin.inds[i] = (
            a.dat_size, a.dat_dim, a.acc, a.ind)
    plan = Plan()
    L = lib()
    L.orc_plan_build.argtypes = [C.POINTER(PlanIn), C.POINTER(Plan)]
    L.orc_plan_check.argtypes = [C.POINTER(PlanIn), C.POINTER(Plan)]
    L.orc_plan_free.argtypes = [C.POINTER(Plan)]
    if L.orc_plan_build(C.byref(pin), C.byref(plan)) != 0:
        raise ValueError("oracle plan build failed")
    try:
        nb, nc, n = plan.nblocks, plan.ncolors, set_size

        def arr(ptr, count, dt=np.int32):
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dt).copy() if count else np.zeros(0, dt)

        out = dict(bsize=plan.bsize, nblocks=nb, ncolors=nc, nshared=plan.nshared,
                   transfer=np.float32(plan.transfer),
                   transfer2=np.float32(plan.transfer2),
                   nthrcol=arr(plan.nthrcol, nb), thrcol=arr(plan.thrcol, n),
                   offset=arr(plan.offset, nb), nelems=arr(plan.nelems, nb),
                   ncolblk=arr(plan.ncolblk, nc), blkmap=arr(plan.blkmap, nb),
                   blkcol=arr(plan.blkcol, nb),
                   ind_offs=arr(plan.ind_offs, nb * ninds),
                   ind_sizes=arr(plan.ind_sizes, nb * ninds),
                   nindirect=np.array(plan.nindirect[:ninds], np.int32),
                   ind_maps=[arr(plan.ind_maps[m], plan.nindirect[m]) for m in range(ninds)],
                   loc_maps=[arr(plan.loc_maps[i], n, np.int16) if args[i].ind >= 0 else None
                             for i in range(len(args))])
        out["check_errors"] = L.orc_plan_check(C.byref(pin), C.byref(plan))
        return out
    finally:
        L.orc_plan_free(C.byref(plan))


def check_plan(set_size: int, args: list, ninds: int, p: dict) -> int:
    """Runs the oracle's plan checks (the reference's op_plan_check,
    common/op_lib_core.c:469-612, plus the exact colouring check) on a plan
    given as arrays (the keys build_plan returns); returns the error count."""
    pin = PlanIn()
    pin.set_size, pin.nargs, pin.ninds, pin.part_size = set_size, len(args), ninds, 0
    keep = []
    for i, a in enumerate(args):
        if a.map is not None:
            keep.append(a.map)
            pin.map[i] = a.map.ctypes.data
        pin.map_dim[i], pin.idx[i], pin.to_size[i] = a.map_dim, a.idx, a.to_size
        pin.dat_size[i], pin.dat_dim[i], pin.acc[i], pin.inds[i] = (
            a.dat_size, a.dat_dim, a.acc, a.ind)
    nb, nc = p["nblocks"], p["ncolors"]
    c_int = lambda v: np.ascontiguousarray(v, np.int32)
    arrays = {k: c_int(p[k]) for k in ("nthrcol", "thrcol", "offset", "nelems", "ncolblk",
                                        "blkmap", "ind_offs", "ind_sizes")}
    # block colour of each block, from blkmap grouped by ncolblk
    blkcol = np.repeat(np.arange(nc, dtype=np.int32), arrays["ncolblk"])
    arrays["blkcol"] = np.zeros(nb, np.int32)
    arrays["blkcol"][arrays["blkmap"]] = blkcol
    keep += list(arrays.values())
    plan = Plan()
    plan.nblocks, plan.ncolors, plan.nshared = nb, nc, p["nshared"]
    plan.ninds, plan.nargs, plan.set_size = ninds, len(args), set_size
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    for k, v in arrays.items():
        setattr(plan, k, ip(v))
    for m in range(ninds):
        im = c_int(p["ind_maps"][m])
        keep.append(im)
        plan.nindirect[m] = im.size
        plan.ind_maps[m] = ip(im)
    for a, lm in enumerate(p["loc_maps"]):
        if lm is not None:
            lm = np.ascontiguousarray(lm, np.int16)
            keep.append(lm)
            plan.loc_maps[a] = lm.ctypes.data_as(C.POINTER(C.c_short))
    L = lib()
    L.orc_plan_check.argtypes = [C.POINTER(PlanIn), C.POINTER(Plan)]
    return int(L.orc_plan_check(C.byref(pin), C.byref(plan)))


def airfoil_plan_args(mesh, loop: str, real_bytes: int) -> tuple:
    """(set_size, args, ninds) of the three indirect airfoil loops, with the
    inds of airfoil/<loop>_kernel.cpp."""
    s = real_bytes
    if loop == "adt_calc":
        args = [PlanArg(mesh.cell, 4, k, mesh.nnode, 2 * s, 2, READ, 0) for k in range(4)]
        args += [PlanArg(dat_size=4 * s, dat_dim=4, acc=READ),
                 PlanArg(dat_size=s, dat_dim=1, acc=WRITE)]
        return mesh.ncell, args, 1
    if loop == "res_calc":
        args = []
        for m, (dim, acc) in enumerate([(2, READ), (4, READ), (1, READ), (4, INC)]):
            mp, to = (mesh.edge, mesh.nnode) if m == 0 else (mesh.ecell, mesh.ncell)
            for k in range(2):
                args.append(PlanArg(mp, 2, k, to, dim * s, dim, acc, m))
        return mesh.nedge, args, 4
    if loop == "bres_calc":
        args = [PlanArg(mesh.bedge, 2, k, mesh.nnode, 2 * s, 2, READ, 0) for k in range(2)]
        for m, (dim, acc) in enumerate([(4, READ), (1, READ), (4, INC)], start=1):
            args.append(PlanArg(mesh.becell.reshape(-1, 1), 1, 0, mesh.ncell, dim * s, dim, acc, m))
        args.append(PlanArg(dat_size=4, dat_dim=1, acc=READ))
        return mesh.nbedge, args, 4
    raise ValueError(loop)


def jacobi_plan_args(prob, real_bytes: int) -> tuple:
    s = real_bytes
    args = [PlanArg(dat_size=s, dat_dim=1, acc=READ),
            PlanArg(prob.pp, 2, 1, prob.nnode, s, 1, READ, 0),
            PlanArg(prob.pp, 2, 0, prob.nnode, s, 1, INC, 1),
            PlanArg(dat_size=0, dat_dim=1, acc=READ)]
    return prob.nedge, args, 2


# ---------------------------------------------------------------------------
# airfoil

class AirfoilState:
    """x, q, qold, adt, res of one form, initialised as
    airfoil/airfoil.cpp:152-157 does (q = qinf, res = 0)."""

    def __init__(self, mesh, consts: np.ndarray, form: str = "f64"):
        real, _ = FORMS[form]
        self.form = form
        self.x = np.ascontiguousarray(mesh.x.astype(real))
        qinf = consts[6:10].astype(real)
        self.q = np.ascontiguousarray(np.tile(qinf, (mesh.ncell, 1)))
        self.qold = np.zeros_like(self.q)
        self.res = np.zeros_like(self.q)
        self.adt = np.zeros(mesh.ncell, dtype=real)


def _fn(name: str, form: str):
    return getattr(lib(), f"orc_{name}_{form}")


def airfoil_seq(mesh, consts, niter: int, form: str = "f64", state=None):
    """Sequential-order time march; returns (state, rms history)."""
    st = state or AirfoilState(mesh, consts, form)
    hist = np.zeros(max(niter, 1))
    c = np.ascontiguousarray(consts, dtype=np.float64)
    rc = _fn("airfoil_seq", form)(
        C.c_int(mesh.nnode), C.c_int(mesh.ncell), C.c_int(mesh.nedge), C.c_int(mesh.nbedge),
        _p(mesh.cell), _p(mesh.edge), _p(mesh.ecell), _p(mesh.bedge), _p(mesh.becell),
        _p(mesh.bound), _p(st.x), _p(st.q), _p(st.qold), _p(st.adt), _p(st.res), _p(c),
        C.c_int(niter), _p(hist))
    if rc:
        raise RuntimeError("oracle airfoil_seq failed")
    return st, hist[:niter]


def airfoil_omp(mesh, consts, niter: int, form: str = "f64", part_size: int = 0,
                nthreads: int = 0, state=None):
    """Reference OpenMP schedule; returns (state, rms history, loop seconds
    [save_soln, adt_calc, res_calc, bres_calc, update])."""
    st = state or AirfoilState(mesh, consts, form)
    hist = np.zeros(max(niter, 1))
    secs = np.zeros(5)
    c = np.ascontiguousarray(consts, dtype=np.float64)
    rc = _fn("airfoil_omp", form)(
        C.c_int(mesh.nnode), C.c_int(mesh.ncell), C.c_int(mesh.nedge), C.c_int(mesh.nbedge),
        _p(mesh.cell), _p(mesh.edge), _p(mesh.ecell), _p(mesh.bedge), _p(mesh.becell),
        _p(mesh.bound), _p(st.x), _p(st.q), _p(st.qold), _p(st.adt), _p(st.res), _p(c),
        C.c_int(niter), _p(hist), C.c_int(part_size), C.c_int(nthreads), _p(secs))
    if rc:
        raise RuntimeError("oracle airfoil_omp failed")
    return st, hist[:niter], secs


def save_soln(form, q, qold):
    _fn("save_soln", form)(C.c_int(q.shape[0]), _p(q), _p(qold))


def adt_calc(form, mesh, x, q, adt, consts):
    _fn("adt_calc", form)(C.c_int(mesh.ncell), _p(mesh.cell), _p(x), _p(q), _p(adt),
                          _p(np.ascontiguousarray(consts, np.float64)))


def res_calc(form, mesh, x, q, adt, res, consts):
    _fn("res_calc", form)(C.c_int(mesh.nedge), _p(mesh.edge), _p(mesh.ecell), _p(x), _p(q),
                          _p(adt), _p(res), _p(np.ascontiguousarray(consts, np.float64)))


def bres_calc(form, mesh, x, q, adt, res, consts):
    _fn("bres_calc", form)(C.c_int(mesh.nbedge), _p(mesh.bedge), _p(mesh.becell),
                           _p(mesh.bound), _p(x), _p(q), _p(adt), _p(res),
                           _p(np.ascontiguousarray(consts, np.float64)))


def update(form, qold, q, res, adt) -> float:
    _, upd = FORMS[form]
    rms = np.zeros(1, dtype=upd)
    _fn("update", form)(C.c_int(q.shape[0]), _p(qold), _p(q), _p(res), _p(adt), _p(rms))
    return float(rms[0])


# ---------------------------------------------------------------------------
# Jacobi

def jacobi_seq(prob, niter: int, alpha: float, beta: float, form: str = "f64"):
    """jac1 iteration in sequential order; returns (u, du, u_sum hist, u_max hist)."""
    real, _ = FORMS[form]
    A = prob.A.astype(real)
    r = prob.r.astype(real)
    u = prob.u.astype(real).copy()
    du = prob.du.astype(real).copy()
    hs = np.zeros(max(niter, 1))
    hm = np.zeros(max(niter, 1))
    f = _fn("jac_seq", form)
    ct = C.c_double if real == np.float64 else C.c_float
    f(C.c_int(prob.nnode), C.c_int(prob.nedge), _p(prob.pp), _p(A), _p(r), _p(u), _p(du),
      ct(alpha), ct(beta), C.c_int(niter), _p(hs), _p(hm))
    return u, du, hs[:niter], hm[:niter]


def jac_update(form, r, du, u, alpha: float):
    """One jac1 update over every node (jac1/update.h, element order):
    u += du + alpha r, du = 0; returns (u_sum, u_max) accumulated in the
    form's reduction type from zero and minus infinity."""
    real, upd = FORMS[form]
    s = np.zeros(1, dtype=upd)
    m = np.full(1, -np.inf, dtype=upd)
    ct = C.c_double if real == np.float64 else C.c_float
    _fn("jac_update", form)(C.c_int(u.shape[0]), _p(r), _p(du), _p(u), _p(s), _p(m), ct(alpha))
    return float(s[0]), float(m[0])


def jacobi_omp(prob, niter: int, alpha: float, beta: float, form: str = "f64",
               part_size: int = 0, nthreads: int = 0):
    real, _ = FORMS[form]
    A = prob.A.astype(real)
    r = prob.r.astype(real)
    u = prob.u.astype(real).copy()
    du = prob.du.astype(real).copy()
    hs = np.zeros(max(niter, 1))
    hm = np.zeros(max(niter, 1))
    f = _fn("jac_omp", form)
    ct = C.c_double if real == np.float64 else C.c_float
    rc = f(C.c_int(prob.nnode), C.c_int(prob.nedge), _p(prob.pp), _p(A), _p(r), _p(u), _p(du),
           ct(alpha), ct(beta), C.c_int(niter), _p(hs), _p(hm), C.c_int(part_size),
           C.c_int(nthreads))
    if rc:
        raise RuntimeError("oracle jac_omp failed")
    return u, du, hs[:niter], hm[:niter]
