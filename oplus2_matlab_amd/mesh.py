"""Unstructured-mesh inputs for the airfoil and Jacobi loops.

The airfoil mesh is a quadrilateral O-mesh around a unit circle: NI cells
around (periodic in the angle) by NJ cells outward, nodes at radius
r_j = 20**(j/NJ) (geometric spacing, near-square cells) and angle
theta_i = 2*pi*i/NI.  The inner circle is a wall (bound 1), the outer circle a
far field (bound 2).

Sets and maps follow the reference airfoil program
(airfoil/airfoil.cpp:165-181): nodes, edges, bedges, cells; pedge (edge->2
nodes), pecell (edge->2 cells), pbedge (bedge->2 nodes), pbecell (bedge->1
cell), pcell (cell->4 nodes).  Orientation follows the reference grid
generator (airfoil/naca0012.m:240-308): with d = x(n1) - x(n2) the face normal
(dy, -dx) points from ecell[0] to ecell[1], and for a boundary edge it points
out of the cell.  Cell nodes go round the cell anticlockwise.

Orderings:
  "cartesian"  cells and nodes numbered i-fastest; each cell is followed by
               its left radial edge and its inner circumferential edge.
  "random"     every set (nodes, cells, edges, bedges) renumbered by a seeded
               permutation (numpy PCG64, the seed is recorded in the mesh).
  "rcm"        reverse Cuthill-McKee order of the cell adjacency graph; nodes
               numbered by first use in cell order, edges by their lower cell.

The text format written by write_grid_dat is the one the reference reads
(airfoil/airfoil.cpp:85-127, written by airfoil/naca0012.m:328-349).
"""

from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass, field

import numpy as np

MESH_MAGIC = 0x4F50324D  # "OP2M"
MESH_VERSION = 1


@dataclass
class AirfoilMesh:
    """Arrays of one airfoil mesh, all C-contiguous (int32 maps, float64 x)."""

    x: np.ndarray        # (nnode, 2) float64
    cell: np.ndarray     # (ncell, 4) int32   pcell
    edge: np.ndarray     # (nedge, 2) int32   pedge
    ecell: np.ndarray    # (nedge, 2) int32   pecell
    bedge: np.ndarray    # (nbedge, 2) int32  pbedge
    becell: np.ndarray   # (nbedge,)  int32   pbecell
    bound: np.ndarray    # (nbedge,)  int32   1 wall, 2 far field
    ni: int = 0
    nj: int = 0
    ordering: str = "cartesian"
    seed: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def nnode(self) -> int:
        return int(self.x.shape[0])

    @property
    def ncell(self) -> int:
        return int(self.cell.shape[0])

    @property
    def nedge(self) -> int:
        return int(self.edge.shape[0])

    @property
    def nbedge(self) -> int:
        return int(self.bedge.shape[0])

    def counts(self):
        return self.nnode, self.ncell, self.nedge, self.nbedge


def o_mesh(ni: int, nj: int, ordering: str = "cartesian", seed: int = 1,
           r_outer: float = 20.0) -> AirfoilMesh:
    """Build an NI x NJ O-mesh in the given ordering (see module docstring)."""
    if ni < 3 or nj < 1:
        raise ValueError("O-mesh needs ni >= 3 and nj >= 1")
    i_n = np.arange(ni, dtype=np.int64)
    j_n = np.arange(nj + 1, dtype=np.int64)
    r = r_outer ** (j_n / nj)
    th = 2.0 * math.pi * i_n / ni
    # node (i, j) -> i + ni*j
    rr, tt = np.meshgrid(r, th, indexing="ij")          # (nj+1, ni)
    x = np.empty(((nj + 1) * ni, 2), dtype=np.float64)
    x[:, 0] = (rr * np.cos(tt)).ravel()
    x[:, 1] = (rr * np.sin(tt)).ravel()

    def node(i, j):
        return (i % ni) + ni * j

    def cellid(i, j):
        return (i % ni) + ni * j

    jj, ii = np.meshgrid(np.arange(nj), np.arange(ni), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    # anticlockwise: (i,j) -> (i,j+1) -> (i+1,j+1) -> (i+1,j)
    cell = np.stack([node(ii, jj), node(ii, jj + 1), node(ii + 1, jj + 1),
                     node(ii + 1, jj)], axis=1)

    # per cell (i,j): radial edge on its low-theta side, then (j>0) the
    # circumferential edge on its inner side.
    # radial edge at theta_i between cell(i-1,j) and cell(i,j):
    #   n1 = node(i,j) (inner), n2 = node(i,j+1); normal points +theta.
    rad_e = np.stack([node(ii, jj), node(ii, jj + 1)], axis=1)
    rad_c = np.stack([cellid(ii - 1, jj), cellid(ii, jj)], axis=1)
    # circumferential edge at r_j between cell(i,j-1) and cell(i,j):
    #   n1 = node(i+1,j), n2 = node(i,j); normal points outward.
    has_c = jj > 0
    cir_e = np.stack([node(ii + 1, jj), node(ii, jj)], axis=1)
    cir_c = np.stack([cellid(ii, jj - 1), cellid(ii, jj)], axis=1)
    nc = ni * nj
    per_cell_e = np.full((nc, 2, 2), -1, dtype=np.int64)
    per_cell_c = np.full((nc, 2, 2), -1, dtype=np.int64)
    per_cell_e[:, 0] = rad_e
    per_cell_c[:, 0] = rad_c
    per_cell_e[:, 1] = cir_e
    per_cell_c[:, 1] = cir_c
    keep = np.stack([np.ones(nc, dtype=bool), has_c], axis=1).ravel()
    edge = per_cell_e.reshape(-1, 2)[keep]
    ecell = per_cell_c.reshape(-1, 2)[keep]

    # boundary edges: wall at j=0 then far field at j=nj, both in i order
    i1 = np.arange(ni)
    wall_e = np.stack([node(i1, 0), node(i1 + 1, 0)], axis=1)
    wall_c = cellid(i1, 0)
    far_e = np.stack([node(i1 + 1, nj), node(i1, nj)], axis=1)
    far_c = cellid(i1, nj - 1)
    bedge = np.concatenate([wall_e, far_e])
    becell = np.concatenate([wall_c, far_c])
    bound = np.concatenate([np.ones(ni, dtype=np.int64),
                            np.full(ni, 2, dtype=np.int64)])

    mesh = AirfoilMesh(
        x=np.ascontiguousarray(x),
        cell=np.ascontiguousarray(cell, dtype=np.int32),
        edge=np.ascontiguousarray(edge, dtype=np.int32),
        ecell=np.ascontiguousarray(ecell, dtype=np.int32),
        bedge=np.ascontiguousarray(bedge, dtype=np.int32),
        becell=np.ascontiguousarray(becell, dtype=np.int32),
        bound=np.ascontiguousarray(bound, dtype=np.int32),
        ni=ni, nj=nj, ordering="cartesian", seed=0)
    if ordering == "cartesian":
        return mesh
    if ordering == "random":
        return permute_random(mesh, seed)
    if ordering == "rcm":
        return renumber_rcm(mesh)
    raise ValueError(f"unknown ordering {ordering!r}")


def apply_permutation(mesh: AirfoilMesh, node_new: np.ndarray,
                      cell_new: np.ndarray, edge_order: np.ndarray,
                      bedge_order: np.ndarray, ordering: str,
                      seed: int = 0) -> AirfoilMesh:
    """Renumber a mesh.

    node_new[old] / cell_new[old] give the new number of each node / cell;
    edge_order / bedge_order list old edge ids in their new order.
    """
    node_new = np.asarray(node_new, dtype=np.int64)
    cell_new = np.asarray(cell_new, dtype=np.int64)
    x = np.empty_like(mesh.x)
    x[node_new] = mesh.x
    cell = np.empty_like(mesh.cell)
    cell[cell_new] = node_new[mesh.cell]
    edge = node_new[mesh.edge[edge_order]]
    ecell = cell_new[mesh.ecell[edge_order]]
    bedge = node_new[mesh.bedge[bedge_order]]
    becell = cell_new[mesh.becell[bedge_order]]
    bound = mesh.bound[bedge_order]
    return AirfoilMesh(
        x=np.ascontiguousarray(x),
        cell=np.ascontiguousarray(cell, dtype=np.int32),
        edge=np.ascontiguousarray(edge, dtype=np.int32),
        ecell=np.ascontiguousarray(ecell, dtype=np.int32),
        bedge=np.ascontiguousarray(bedge, dtype=np.int32),
        becell=np.ascontiguousarray(becell, dtype=np.int32),
        bound=np.ascontiguousarray(bound, dtype=np.int32),
        ni=mesh.ni, nj=mesh.nj, ordering=ordering, seed=seed)


def permute_random(mesh: AirfoilMesh, seed: int) -> AirfoilMesh:
    """Seeded random renumbering of all four sets (numpy PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    node_new = rng.permutation(mesh.nnode)
    cell_new = rng.permutation(mesh.ncell)
    edge_order = rng.permutation(mesh.nedge)
    bedge_order = rng.permutation(mesh.nbedge)
    return apply_permutation(mesh, node_new, cell_new, edge_order,
                             bedge_order, "random", seed)


def renumber_rcm(mesh: AirfoilMesh) -> AirfoilMesh:
    """Reverse Cuthill-McKee renumbering of cells; nodes and edges follow."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    nc = mesh.ncell
    a = mesh.ecell[:, 0].astype(np.int64)
    b = mesh.ecell[:, 1].astype(np.int64)
    g = coo_matrix((np.ones(2 * a.size, dtype=np.int8),
                    (np.concatenate([a, b]), np.concatenate([b, a]))),
                   shape=(nc, nc)).tocsr()
    order = np.asarray(reverse_cuthill_mckee(g, symmetric_mode=True),
                       dtype=np.int64)
    cell_new = np.empty(nc, dtype=np.int64)
    cell_new[order] = np.arange(nc)
    # nodes: first use while walking cells in their new order
    walk = mesh.cell[order].ravel()
    _, first = np.unique(walk, return_index=True)
    node_rank = np.empty(mesh.nnode, dtype=np.int64)
    used = walk[np.sort(first)]
    node_rank[used] = np.arange(used.size)
    # edges: by the lower of their two new cell numbers, then the higher
    ec = cell_new[mesh.ecell]
    lo = ec.min(axis=1)
    hi = ec.max(axis=1)
    edge_order = np.lexsort((hi, lo))
    bedge_order = np.argsort(cell_new[mesh.becell], kind="stable")
    return apply_permutation(mesh, node_rank, cell_new, edge_order,
                             bedge_order, "rcm")


# ---------------------------------------------------------------------------
# file formats

def write_grid_dat(mesh: AirfoilMesh, path: str) -> None:
    """Write the reference text format (airfoil/airfoil.cpp:85-127).

    The reference reads coordinates into floats, so they are written as the
    float32 values, with enough digits to round-trip exactly."""
    with open(path, "w") as f:
        f.write(" %d %d %d %d \n" % mesh.counts())
        np.savetxt(f, mesh.x.astype(np.float32).astype(np.float64), fmt=" %.9g %.9g ")
        np.savetxt(f, mesh.cell, fmt=" %d %d %d %d ")
        np.savetxt(f, np.concatenate([mesh.edge, mesh.ecell], axis=1),
                   fmt=" %d %d %d %d ")
        np.savetxt(f, np.stack([mesh.bedge[:, 0], mesh.bedge[:, 1],
                                mesh.becell, mesh.bound], axis=1),
                   fmt=" %d %d %d %d ")


def read_grid_dat(path: str) -> AirfoilMesh:
    """Read the reference text format (airfoil/airfoil.cpp:85-127)."""
    with open(path) as f:
        nnode, ncell, nedge, nbedge = (int(v) for v in f.readline().split())
        x = np.loadtxt(f, max_rows=nnode, dtype=np.float64, ndmin=2)
        cell = np.loadtxt(f, max_rows=ncell, dtype=np.int32, ndmin=2)
        e = np.loadtxt(f, max_rows=nedge, dtype=np.int32, ndmin=2)
        b = np.loadtxt(f, max_rows=nbedge, dtype=np.int32, ndmin=2)
    return AirfoilMesh(
        x=np.ascontiguousarray(x), cell=np.ascontiguousarray(cell),
        edge=np.ascontiguousarray(e[:, :2]), ecell=np.ascontiguousarray(e[:, 2:]),
        bedge=np.ascontiguousarray(b[:, :2]), becell=np.ascontiguousarray(b[:, 2]),
        bound=np.ascontiguousarray(b[:, 3]), ordering="file")


def write_mesh_bin(mesh: AirfoilMesh, path: str) -> None:
    """Binary mesh: int32 header[8] then x (f64), pcell, pedge, pecell,
    pbedge, pbecell, bound (all int32), little endian."""
    hdr = struct.pack("<8i", MESH_MAGIC, MESH_VERSION, mesh.nnode, mesh.ncell,
                      mesh.nedge, mesh.nbedge, mesh.ni, mesh.nj)
    with open(path, "wb") as f:
        f.write(hdr)
        for a in (mesh.x.astype("<f8"), mesh.cell, mesh.edge, mesh.ecell,
                  mesh.bedge, mesh.becell, mesh.bound):
            f.write(np.ascontiguousarray(a).astype(
                a.dtype.newbyteorder("<")).tobytes())


def read_mesh_bin(path: str) -> AirfoilMesh:
    with open(path, "rb") as f:
        magic, ver, nn, nc, ne, nb, ni, nj = struct.unpack("<8i", f.read(32))
        if magic != MESH_MAGIC or ver != MESH_VERSION:
            raise ValueError(f"{path}: not a mesh file")

        def take(n, dt):
            return np.frombuffer(f.read(n * np.dtype(dt).itemsize), dtype=dt).copy()
        x = take(2 * nn, "<f8").reshape(nn, 2)
        cell = take(4 * nc, "<i4").reshape(nc, 4)
        edge = take(2 * ne, "<i4").reshape(ne, 2)
        ecell = take(2 * ne, "<i4").reshape(ne, 2)
        bedge = take(2 * nb, "<i4").reshape(nb, 2)
        becell = take(nb, "<i4")
        bound = take(nb, "<i4")
    return AirfoilMesh(x=x, cell=cell, edge=edge, ecell=ecell, bedge=bedge,
                       becell=becell, bound=bound, ni=ni, nj=nj,
                       ordering="file")


# ---------------------------------------------------------------------------
# flow constants

@dataclass(frozen=True)
class FlowConstants:
    """Free-stream constants, computed as airfoil/airfoil.cpp:135-150 does,
    with the flow direction alpha applied to qinf (the reference computes
    alpha but leaves qinf[2] = 0; alpha=0 reproduces it exactly)."""

    gam: float
    gm1: float
    cfl: float
    eps: float
    mach: float
    alpha: float
    qinf: tuple

    def as_array(self) -> np.ndarray:
        """[gam, gm1, cfl, eps, mach, alpha, qinf0..3] as float64."""
        return np.array([self.gam, self.gm1, self.cfl, self.eps, self.mach,
                         self.alpha, *self.qinf], dtype=np.float64)


def flow_constants(mach: float = 0.4, alpha_deg: float = 3.0,
                   gam: float = 1.4, cfl: float = 0.9,
                   eps: float = 0.05) -> FlowConstants:
    gm1 = gam - 1.0
    alpha = alpha_deg * math.atan(1.0) / 45.0
    p = 1.0
    r = 1.0
    u = math.sqrt(gam * p / r) * mach
    e = p / (r * gm1) + 0.5 * u * u
    qinf = (r, r * u * math.cos(alpha), r * u * math.sin(alpha), r * e)
    return FlowConstants(gam, gm1, cfl, eps, mach, alpha, qinf)


def flow_constants_f32_reference() -> FlowConstants:
    """The reference's single-precision constants, bit for bit
    (airfoil/airfoil.cpp:135-150: float arithmetic, qinf[2] = 0)."""
    f = np.float32
    gam = f(1.4)
    gm1 = f(gam - f(1.0))
    cfl = f(0.9)
    eps = f(0.05)
    mach = f(0.4)
    alpha = f(f(3.0) * f(math.atan(f(1.0))) / f(45.0))
    p = f(1.0)
    r = f(1.0)
    u = f(f(np.sqrt(f(gam * p / r))) * mach)
    e = f(p / f(r * gm1) + f(f(f(0.5) * u) * u))
    qinf = (float(r), float(f(r * u)), 0.0, float(f(r * e)))
    return FlowConstants(float(gam), float(gm1), float(cfl), float(eps),
                         float(mach), float(alpha), qinf)


# ---------------------------------------------------------------------------
# Jacobi problem

@dataclass
class JacobiProblem:
    """5-point Jacobi on an nn x nn node grid (the jac1 program's problem,
    jac1/jac.cpp:74-120, with the self and neighbour weights as given)."""

    pp: np.ndarray   # (nedge, 2) int32: pp[e,0] = du target, pp[e,1] = u source
    A: np.ndarray    # (nedge,) float64
    r: np.ndarray    # (nnode,) float64
    u: np.ndarray    # (nnode,) float64
    du: np.ndarray   # (nnode,) float64
    nn: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def nnode(self) -> int:
        return int(self.r.shape[0])

    @property
    def nedge(self) -> int:
        return int(self.A.shape[0])


def jacobi_grid(nn: int = 64, seed: int = 1, a_self: float = 4.0,
                a_nbr: float = -1.0, permute: bool = True) -> JacobiProblem:
    """nn*nn nodes; per node a self edge then one edge per grid neighbour.

    r is uniform in [0,1) from the seed (numpy PCG64) and u = du = 0.  With
    permute, nodes and edges are renumbered by seeded permutations.
    """
    n = np.arange(nn * nn, dtype=np.int64)
    i = n % nn
    j = n // nn
    pp_l, a_l = [], []
    for node in range(nn * nn):
        ii, jj = node % nn, node // nn
        pp_l.append((node, node))
        a_l.append(a_self)
        for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            i2, j2 = ii + di, jj + dj
            if 0 <= i2 < nn and 0 <= j2 < nn:
                pp_l.append((node, i2 + nn * j2))
                a_l.append(a_nbr)
    del i, j
    pp = np.array(pp_l, dtype=np.int64)
    A = np.array(a_l, dtype=np.float64)
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.random(nn * nn)
    if permute:
        node_new = rng.permutation(nn * nn)
        edge_order = rng.permutation(pp.shape[0])
        r2 = np.empty_like(r)
        r2[node_new] = r
        r = r2
        pp = node_new[pp[edge_order]]
        A = A[edge_order]
    return JacobiProblem(pp=np.ascontiguousarray(pp, dtype=np.int32),
                         A=np.ascontiguousarray(A), r=np.ascontiguousarray(r),
                         u=np.zeros(nn * nn), du=np.zeros(nn * nn), nn=nn)


def write_jacobi_bin(p: JacobiProblem, path: str) -> None:
    """int32 header[4] = (magic, version, nnode, nedge), then pp (int32),
    A, r, u, du (float64)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<4i", MESH_MAGIC + 1, MESH_VERSION, p.nnode, p.nedge))
        for a in (p.pp.astype("<i4"), p.A.astype("<f8"), p.r.astype("<f8"),
                  p.u.astype("<f8"), p.du.astype("<f8")):
            f.write(a.tobytes())


def jacobi_reference_problem(nn: int = 6) -> JacobiProblem:
    """The jac1 program's own problem (jac1/jac.cpp:74-120): (nn-1)^2 interior
    nodes, a self edge with A = -1 then one edge of weight 0.25 per interior
    neighbour, r = 0.25 per missing (boundary) neighbour; run with alpha =
    beta = 1 (jac1/jac.cpp:138,145)."""
    m = nn - 1
    pp_l, a_l = [], []
    r = np.zeros(m * m)
    for i in range(1, nn):
        for j in range(1, nn):
            n = i - 1 + (j - 1) * m
            pp_l.append((n, n))
            a_l.append(-1.0)
            for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                i2, j2 = i + di, j + dj
                if i2 == 0 or i2 == nn or j2 == 0 or j2 == nn:
                    r[n] += 0.25
                else:
                    pp_l.append((n, i2 - 1 + (j2 - 1) * m))
                    a_l.append(0.25)
    return JacobiProblem(pp=np.array(pp_l, dtype=np.int32), A=np.array(a_l),
                         r=r, u=np.zeros(m * m), du=np.zeros(m * m), nn=nn)


# ---------------------------------------------------------------------------
# batches of meshes

@dataclass
class MeshBatch:
    """Several airfoil meshes as one union mesh: node, cell, edge and bedge
    numbers of mesh i are offset by the sizes of meshes 0..i-1, so one set of
    each kind covers the batch.  cmesh / bmesh map cells / bedges to their
    mesh; qinf holds each mesh's free stream (4 values per mesh)."""

    union: AirfoilMesh
    meshes: list
    consts: list            # flow-constant arrays, one per mesh
    cmesh: np.ndarray       # (ncell,) int32
    bmesh: np.ndarray       # (nbedge,) int32
    qinf: np.ndarray        # (nmesh, 4) float64
    cell_off: np.ndarray    # (nmesh + 1,)

    @property
    def nmesh(self) -> int:
        return len(self.meshes)


def mesh_batch(nmesh: int = 64, seed: int = 1, cells_lo: int = 10_000,
               cells_hi: int = 400_000, ordering: str = "cartesian") -> MeshBatch:
    """nmesh O-meshes with cell counts log-uniform in [cells_lo, cells_hi]
    (NI = 2 NJ), mach uniform in [0.3, 0.6] and alpha uniform in [0, 5]
    degrees, all from `seed` (numpy PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cells = np.exp(rng.uniform(np.log(cells_lo), np.log(cells_hi), nmesh))
    mach = rng.uniform(0.3, 0.6, nmesh)
    alpha = rng.uniform(0.0, 5.0, nmesh)
    meshes, consts = [], []
    for i in range(nmesh):
        nj = max(2, int(round(math.sqrt(cells[i] / 2.0))))
        meshes.append(o_mesh(2 * nj, nj, ordering, seed=seed * 1000 + i))
        consts.append(flow_constants(mach=float(mach[i]), alpha_deg=float(alpha[i])).as_array())
    return concat_meshes(meshes, consts)


def concat_meshes(meshes: list, consts: list) -> MeshBatch:
    nn = np.cumsum([0] + [m.nnode for m in meshes])
    nc = np.cumsum([0] + [m.ncell for m in meshes])
    ne = np.cumsum([0] + [m.nedge for m in meshes])
    nb = np.cumsum([0] + [m.nbedge for m in meshes])
    union = AirfoilMesh(
        x=np.concatenate([m.x for m in meshes]),
        cell=np.concatenate([m.cell + nn[i] for i, m in enumerate(meshes)]).astype(np.int32),
        edge=np.concatenate([m.edge + nn[i] for i, m in enumerate(meshes)]).astype(np.int32),
        ecell=np.concatenate([m.ecell + nc[i] for i, m in enumerate(meshes)]).astype(np.int32),
        bedge=np.concatenate([m.bedge + nn[i] for i, m in enumerate(meshes)]).astype(np.int32),
        becell=np.concatenate([m.becell + nc[i] for i, m in enumerate(meshes)]).astype(np.int32),
        bound=np.concatenate([m.bound for m in meshes]).astype(np.int32),
        ordering="batch")
    cmesh = np.repeat(np.arange(len(meshes), dtype=np.int32), [m.ncell for m in meshes])
    bmesh = np.repeat(np.arange(len(meshes), dtype=np.int32), [m.nbedge for m in meshes])
    qinf = np.stack([c[6:10] for c in consts])
    del ne, nb
    return MeshBatch(union=union, meshes=meshes, consts=consts, cmesh=cmesh, bmesh=bmesh,
                     qinf=qinf, cell_off=nc)
