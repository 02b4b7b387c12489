"""The airfoil and Jacobi programs over the OP2 API, run on the GPU.

`run_airfoil` is the time-march of the reference airfoil program
(airfoil/airfoil.cpp:152-257) written against op2.py; `run_jacobi` is the
iteration of jac1/jac.cpp:122-169.  Both declare everything through the
public API, so they exercise the same path a C/C++ program takes.
"""

from __future__ import annotations

import numpy as np

from . import op2
from .op2 import (OP_ID, OP_INC, OP_MAX, OP_READ, OP_RW, OP_WRITE, op_arg_dat, op_arg_gbl,
                  op_decl_const2, op_decl_dat, op_decl_map, op_decl_set, op_par_loop)


class Airfoil:
    """Sets, maps and dats of one airfoil mesh, declared as
    airfoil/airfoil.cpp:165-189 declares them."""

    def __init__(self, mesh, consts: np.ndarray, dtype: str = "double", q0=None, res0=None,
                 renumber: bool = False):
        self.real = np.float64 if dtype == "double" else np.float32
        self.t = dtype
        self.mesh = mesh
        self.ncell = mesh.ncell
        real = self.real
        self.x = np.ascontiguousarray(mesh.x.astype(real).reshape(-1))
        qinf = np.asarray(consts[6:10], dtype=real)
        self.q = np.ascontiguousarray(np.tile(qinf, mesh.ncell))
        if q0 is not None:
            self.q[:] = np.asarray(q0, dtype=real).reshape(-1)
        self.qold = np.zeros_like(self.q)
        self.res = np.zeros_like(self.q)
        if res0 is not None:
            self.res[:] = np.asarray(res0, dtype=real).reshape(-1)
        self.adt = np.zeros(mesh.ncell, dtype=real)
        self._graph_iters = {}  # graph id -> iterations it holds
        self._single = None     # a one-iteration graph for remainders

        self.nodes = op_decl_set(mesh.nnode, "nodes")
        self.edges = op_decl_set(mesh.nedge, "edges")
        self.bedges = op_decl_set(mesh.nbedge, "bedges")
        self.cells = op_decl_set(mesh.ncell, "cells")
        self.pedge = op_decl_map(self.edges, self.nodes, 2, mesh.edge, "pedge")
        self.pecell = op_decl_map(self.edges, self.cells, 2, mesh.ecell, "pecell")
        self.pbedge = op_decl_map(self.bedges, self.nodes, 2, mesh.bedge, "pbedge")
        self.pbecell = op_decl_map(self.bedges, self.cells, 1, mesh.becell, "pbecell")
        self.pcell = op_decl_map(self.cells, self.nodes, 4, mesh.cell, "pcell")
        self.p_bound = op_decl_dat(self.bedges, 1, "int", np.ascontiguousarray(mesh.bound), "p_bound")
        self.p_x = op_decl_dat(self.nodes, 2, dtype, self.x, "p_x")
        self.p_q = op_decl_dat(self.cells, 4, dtype, self.q, "p_q")
        self.p_qold = op_decl_dat(self.cells, 4, dtype, self.qold, "p_qold")
        self.p_adt = op_decl_dat(self.cells, 1, dtype, self.adt, "p_adt")
        self.p_res = op_decl_dat(self.cells, 4, dtype, self.res, "p_res")
        names = ("gam", "gm1", "cfl", "eps", "mach", "alpha")
        for i, n in enumerate(names):
            op_decl_const2(n, 1, dtype, np.array([consts[i]], dtype=real))
        op_decl_const2("qinf", 4, dtype, np.asarray(consts[6:10], dtype=real))
        self.rms = np.zeros(1, dtype=np.float64)
        if renumber:
            # the runtime reorders cells (and through the maps edges, nodes,
            # bedges) for locality; fetched data keeps the mesh's numbering
            op2.renumber(self.pecell)

    # the five loops, with the arguments of airfoil/airfoil.cpp:201-248
    def save_soln(self):
        t = self.t
        op_par_loop(op2.save_soln, "save_soln", self.cells,
                    op_arg_dat(self.p_q, -1, OP_ID, 4, t, OP_READ),
                    op_arg_dat(self.p_qold, -1, OP_ID, 4, t, OP_WRITE))

    def adt_calc(self):
        t = self.t
        op_par_loop(op2.adt_calc, "adt_calc", self.cells,
                    op_arg_dat(self.p_x, 0, self.pcell, 2, t, OP_READ),
                    op_arg_dat(self.p_x, 1, self.pcell, 2, t, OP_READ),
                    op_arg_dat(self.p_x, 2, self.pcell, 2, t, OP_READ),
                    op_arg_dat(self.p_x, 3, self.pcell, 2, t, OP_READ),
                    op_arg_dat(self.p_q, -1, OP_ID, 4, t, OP_READ),
                    op_arg_dat(self.p_adt, -1, OP_ID, 1, t, OP_WRITE))

    def res_calc_args(self):
        t = self.t
        return (op_arg_dat(self.p_x, 0, self.pedge, 2, t, OP_READ),
                op_arg_dat(self.p_x, 1, self.pedge, 2, t, OP_READ),
                op_arg_dat(self.p_q, 0, self.pecell, 4, t, OP_READ),
                op_arg_dat(self.p_q, 1, self.pecell, 4, t, OP_READ),
                op_arg_dat(self.p_adt, 0, self.pecell, 1, t, OP_READ),
                op_arg_dat(self.p_adt, 1, self.pecell, 1, t, OP_READ),
                op_arg_dat(self.p_res, 0, self.pecell, 4, t, OP_INC),
                op_arg_dat(self.p_res, 1, self.pecell, 4, t, OP_INC))

    def bres_calc_args(self):
        t = self.t
        return (op_arg_dat(self.p_x, 0, self.pbedge, 2, t, OP_READ),
                op_arg_dat(self.p_x, 1, self.pbedge, 2, t, OP_READ),
                op_arg_dat(self.p_q, 0, self.pbecell, 4, t, OP_READ),
                op_arg_dat(self.p_adt, 0, self.pbecell, 1, t, OP_READ),
                op_arg_dat(self.p_res, 0, self.pbecell, 4, t, OP_INC),
                op_arg_dat(self.p_bound, -1, OP_ID, 1, "int", OP_READ))

    def autotune(self, reps: int = 5, online: bool = False) -> dict:
        """Picks strategy and part size of res_calc and bres_calc on this
        mesh.  Offline: the runtime times each loop alone on the live data
        (state unchanged).  Online: the loops are timed inside real
        iterations, which advance the solution; returns the choices and the
        number of iterations that took."""
        if not online:
            self.adt_calc()
            return {"res_calc": op2.autotune("res_calc", self.edges, self.res_calc_args(), reps),
                    "bres_calc": op2.autotune("bres_calc", self.bedges, self.bres_calc_args(),
                                              reps)}
        names = ("res_calc", "bres_calc")
        for n in names:
            op2.autotune_online(n, reps)
        for it in range(1, 1000):
            self.iterate()
            got = {n: op2.autotune_online_result(n) for n in names}
            if all(got.values()):
                got["iterations"] = it
                return got
        raise op2.OP2Error("online autotuning did not finish")

    def res_calc(self):
        op_par_loop(op2.res_calc, "res_calc", self.edges, *self.res_calc_args())

    def bres_calc(self):
        op_par_loop(op2.bres_calc, "bres_calc", self.bedges, *self.bres_calc_args())

    def update(self) -> float:
        t = self.t
        self.rms[0] = 0.0
        op_par_loop(op2.update, "update", self.cells,
                    op_arg_dat(self.p_qold, -1, OP_ID, 4, t, OP_READ),
                    op_arg_dat(self.p_q, -1, OP_ID, 4, t, OP_WRITE),
                    op_arg_dat(self.p_res, -1, OP_ID, 4, t, OP_RW),
                    op_arg_dat(self.p_adt, -1, OP_ID, 1, t, OP_READ),
                    op_arg_gbl(self.rms, 1, "double", OP_INC))
        return float(self.rms[0])

    def iterate(self) -> float:
        """One iteration of airfoil/airfoil.cpp:197-257; returns
        sqrt(rms / ncell) of the second stage."""
        self.save_soln()
        rms = 0.0
        for _ in range(2):
            self.adt_calc()
            self.res_calc()
            self.bres_calc()
            rms = self.update()
        return float(np.sqrt(rms / self.ncell))

    def capture_iteration(self, iters: int = 1) -> int:
        """Captures `iters` iterations as one hipGraph (deferred reductions
        on; the loops must have run once before).  Returns the graph id.
        With more than one iteration the runtime also fuses each update with
        the next iteration's save_soln and adt_calc (graph loop fusion)."""
        if iters < 1:
            raise ValueError("capture at least one iteration")
        op2.set_deferred_reductions(True)
        op2.graph_begin()
        try:
            for _ in range(iters):
                self.save_soln()
                for _ in range(2):
                    self.adt_calc()
                    self.res_calc()
                    self.bres_calc()
                    self.update()
        except Exception:
            op2.graph_end()
            raise
        gid = op2.graph_end()
        self._graph_iters[gid] = iters
        return gid

    def _replays(self, gid: int, niter: int):
        """(graph, launches) pairs that run niter iterations: whole replays of
        gid, then a one-iteration graph (captured on first need) for the
        rest."""
        k = self._graph_iters.get(gid, 1)
        out = [(gid, niter // k)] if niter // k else []
        if niter % k:
            if self._single is None:
                self._single = self.capture_iteration(1)
            out.append((self._single, niter % k))
        return out

    def run_graph(self, gid: int, niter: int, chunk: int = 8192) -> np.ndarray:
        """Runs niter iterations by replaying the captured graph; returns
        the rms history from the reduction log (two updates per iteration;
        the second stage's sum gives the iteration's rms)."""
        op2.reduction_log()  # drop earlier entries
        hist = []
        for g, count in self._replays(gid, niter):
            k = self._graph_iters[g]
            per = max(1, chunk // k)
            done = 0
            while done < count:
                n = min(per, count - done)
                op2.graph_launch(g, n)
                log = op2.reduction_log()
                if log.size != 2 * n * k:
                    raise op2.OP2Error(f"reduction log holds {log.size} entries, "
                                       f"expected {2 * n * k}")
                hist.append(np.sqrt(log[1::2] / self.ncell))
                done += n
        return np.concatenate(hist) if hist else np.zeros(0)

    def fetch(self):
        for d in (self.p_q, self.p_qold, self.p_adt, self.p_res):
            op2.op_fetch_data(d)
        return self.q.reshape(-1, 4), self.qold.reshape(-1, 4), self.adt, self.res.reshape(-1, 4)


def run_airfoil(mesh, consts, niter: int, dtype: str = "double", graph: bool = False,
                renumber: bool = False, graph_iters: int = 1):
    """Declares the mesh, runs `niter` iterations, returns (rms history,
    q, adt, res) after fetching the dats.  With graph=True the first
    iteration runs eagerly and the rest replay a captured hipGraph of
    graph_iters iterations with deferred reductions.  The caller owns
    op_init/op_exit."""
    a = Airfoil(mesh, consts, dtype, renumber=renumber)
    if not graph or niter < 2:
        hist = np.array([a.iterate() for _ in range(niter)])
    else:
        first = a.iterate()
        gid = a.capture_iteration(graph_iters)
        hist = np.concatenate([[first], a.run_graph(gid, niter - 1)])
        op2.set_deferred_reductions(False)
    q, _, adt, res = a.fetch()
    return hist, q.copy(), adt.copy(), res.copy()


class AirfoilBatch(Airfoil):
    """A batch of airfoil meshes (mesh.MeshBatch) advanced together as one
    union mesh.  The loops are the airfoil loops with two extra arguments:
    bres_calc reads its boundary edge's mesh's free stream through
    bedge -> mesh (a dat on the meshes set instead of the qinf constant) and
    update increments each mesh's own rms through cell -> mesh.  Every mesh
    then marches exactly as if it ran alone (no edge crosses meshes).

    The per-mesh rms sums come from the runtime's reduction log (deferred
    reductions are switched on): each update call logs one sum per mesh, in
    mesh order."""

    def __init__(self, batch, dtype: str = "double", renumber: bool = False):
        c0 = np.asarray(batch.consts[0], dtype=np.float64)
        for c in batch.consts:
            if not np.array_equal(np.asarray(c)[:4], c0[:4]):
                raise ValueError("meshes of a batch share gam, gm1, cfl and eps")
        real = np.float64 if dtype == "double" else np.float32
        super().__init__(batch.union, c0, dtype, q0=batch.qinf[batch.cmesh].astype(real),
                         renumber=renumber)
        self.batch = batch
        self.nmesh = batch.nmesh
        self.mesh_cells = np.diff(batch.cell_off).astype(np.float64)
        self.meshes = op_decl_set(batch.nmesh, "meshes")
        self.pcmesh = op_decl_map(self.cells, self.meshes, 1, batch.cmesh, "pcmesh")
        self.pbmesh = op_decl_map(self.bedges, self.meshes, 1, batch.bmesh, "pbmesh")
        self.qinf = np.ascontiguousarray(batch.qinf.astype(real).reshape(-1))
        self.p_qinf = op_decl_dat(self.meshes, 4, dtype, self.qinf, "p_qinf")
        self.rms_m = np.zeros(batch.nmesh, dtype=np.float64)
        self.p_rms = op_decl_dat(self.meshes, 1, "double", self.rms_m, "p_rms")
        op2.set_deferred_reductions(True)

    def bres_calc_args(self):
        return super().bres_calc_args() + (
            op_arg_dat(self.p_qinf, 0, self.pbmesh, 4, self.t, OP_READ),)

    def update(self):
        t = self.t
        op_par_loop(op2.update, "update", self.cells,
                    op_arg_dat(self.p_qold, -1, OP_ID, 4, t, OP_READ),
                    op_arg_dat(self.p_q, -1, OP_ID, 4, t, OP_WRITE),
                    op_arg_dat(self.p_res, -1, OP_ID, 4, t, OP_RW),
                    op_arg_dat(self.p_adt, -1, OP_ID, 1, t, OP_READ),
                    op_arg_dat(self.p_rms, 0, self.pcmesh, 1, "double", OP_INC))

    def _history(self, log: np.ndarray, n: int) -> np.ndarray:
        if log.size != 2 * self.nmesh * n:
            raise op2.OP2Error(f"reduction log holds {log.size} entries, "
                               f"expected {2 * self.nmesh * n}")
        second = log.reshape(n, 2, self.nmesh)[:, 1, :]
        return np.sqrt(second / self.mesh_cells)

    def iterate(self) -> np.ndarray:
        """One iteration of every mesh; returns the nmesh values of
        sqrt(rms / ncell) of the second stage."""
        op2.reduction_log()
        self.save_soln()
        for _ in range(2):
            self.adt_calc()
            self.res_calc()
            self.bres_calc()
            self.update()
        return self._history(op2.reduction_log(), 1)[0]

    def run_graph(self, gid: int, niter: int, chunk: int = 0) -> np.ndarray:
        """Runs niter iterations by replaying the captured graph; returns
        the (niter, nmesh) rms history."""
        cap = max(1, op2.REDUCTION_LOG_CAPACITY // (2 * self.nmesh))
        op2.reduction_log()
        hist = []
        for g, count in self._replays(gid, niter):
            k = self._graph_iters[g]
            per = max(1, (min(chunk, cap) if chunk > 0 else cap) // k)
            done = 0
            while done < count:
                n = min(per, count - done)
                op2.graph_launch(g, n)
                hist.append(self._history(op2.reduction_log(), n * k))
                done += n
        return np.concatenate(hist) if hist else np.zeros((0, self.nmesh))

    def mesh_slice(self, arr: np.ndarray, i: int) -> np.ndarray:
        """Rows of mesh i in a per-cell array of the union."""
        off = self.batch.cell_off
        return arr[off[i]:off[i + 1]]


def run_airfoil_batch(batch, niter: int, dtype: str = "double", graph: bool = False,
                      renumber: bool = False, graph_iters: int = 1):
    """Runs `niter` iterations of every mesh of the batch; returns the
    (niter, nmesh) rms history and the union's q after fetching.  The caller
    owns op_init/op_exit."""
    a = AirfoilBatch(batch, dtype, renumber=renumber)
    if not graph or niter < 2:
        hist = np.stack([a.iterate() for _ in range(niter)]) if niter else np.zeros((0, a.nmesh))
    else:
        first = a.iterate()
        gid = a.capture_iteration(graph_iters)
        hist = np.concatenate([first[None, :], a.run_graph(gid, niter - 1)])
    op2.set_deferred_reductions(False)
    q, _, _, _ = a.fetch()
    return hist, q.copy()


class Jacobi:
    """The jac1 program's sets, map and dats (jac1/jac.cpp:128-139)."""

    def __init__(self, prob, alpha: float, dtype: str = "double", renumber: bool = False):
        real = np.float64 if dtype == "double" else np.float32
        self.t = dtype
        self.real = real
        self.A = np.ascontiguousarray(prob.A.astype(real))
        self.r = np.ascontiguousarray(prob.r.astype(real))
        self.u = np.ascontiguousarray(prob.u.astype(real))
        self.du = np.ascontiguousarray(prob.du.astype(real))
        self.nnode = prob.nnode
        self.nodes = op_decl_set(prob.nnode, "nodes")
        self.edges = op_decl_set(prob.nedge, "edges")
        self.ppedge = op_decl_map(self.edges, self.nodes, 2, prob.pp, "ppedge")
        self.p_A = op_decl_dat(self.edges, 1, dtype, self.A, "p_A")
        self.p_r = op_decl_dat(self.nodes, 1, dtype, self.r, "p_r")
        self.p_u = op_decl_dat(self.nodes, 1, dtype, self.u, "p_u")
        self.p_du = op_decl_dat(self.nodes, 1, dtype, self.du, "p_du")
        op_decl_const2("alpha", 1, dtype, np.array([alpha], dtype=real))
        if renumber:
            op2.renumber(self.ppedge)

    def iterate(self, beta: float):
        """One iteration of jac1/jac.cpp:147-163; returns (u_sum, u_max)."""
        t = self.t
        b = np.array([beta], dtype=self.real)
        op_par_loop(op2.res, "res", self.edges,
                    op_arg_dat(self.p_A, -1, OP_ID, 1, t, OP_READ),
                    op_arg_dat(self.p_u, 1, self.ppedge, 1, t, OP_READ),
                    op_arg_dat(self.p_du, 0, self.ppedge, 1, t, OP_INC),
                    op_arg_gbl(b, 1, t, OP_READ))
        u_sum = np.zeros(1, dtype=np.float64)
        u_max = np.zeros(1, dtype=np.float64)
        op_par_loop(op2.update, "update", self.nodes,
                    op_arg_dat(self.p_r, -1, OP_ID, 1, t, OP_READ),
                    op_arg_dat(self.p_du, -1, OP_ID, 1, t, OP_RW),
                    op_arg_dat(self.p_u, -1, OP_ID, 1, t, OP_INC),
                    op_arg_gbl(u_sum, 1, "double", OP_INC),
                    op_arg_gbl(u_max, 1, "double", OP_MAX))
        return float(u_sum[0]), float(u_max[0])


def run_jacobi(prob, niter: int, alpha: float, beta: float, dtype: str = "double",
               renumber: bool = False):
    j = Jacobi(prob, alpha, dtype, renumber=renumber)
    hs, hm = [], []
    for _ in range(niter):
        s, m = j.iterate(beta)
        hs.append(s)
        hm.append(m)
    op2.op_fetch_data(j.p_u)
    op2.op_fetch_data(j.p_du)
    return np.array(hs), np.array(hm), j.u.copy(), j.du.copy()
