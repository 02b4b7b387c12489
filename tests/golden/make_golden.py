"""Writes the golden fixtures under tests/golden/ from the reference itself.

Runs in the development container only (it needs the reference sources):
builds the reference's CPU paths with oracle/ref/Makefile into oracle/_ref/,
runs them on seeded meshes, and stores their outputs as small .npz / .txt
fixtures.  The reference has no fixtures of its own (no test suite, no
new_grid.dat), so these are the only pins that come from the reference.

    python tests/golden/make_golden.py            # all fixtures
    python tests/golden/make_golden.py --quick    # skip the 180k-cell run
"""

from __future__ import annotations

import argparse
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oplus2_matlab_amd import mesh as M  # noqa: E402

GOLD = os.path.dirname(os.path.abspath(__file__))
REFBIN = os.path.join(ROOT, "oracle", "_ref")


def build_ref():
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "oracle", "ref")], check=True)


def run(args, threads=None, cwd=None):
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    return subprocess.run(args, check=True, env=env, cwd=cwd, capture_output=True,
                          text=True).stdout


def read_airfoil_out(path, niter, ncell):
    raw = open(path, "rb").read()
    a = np.frombuffer(raw[16:], np.float64)
    o = 0
    out = {"rms": a[o:o + niter].copy()}
    o += niter
    for name, w in (("q", 4), ("qold", 4), ("adt", 1), ("res", 4)):
        out[name] = a[o:o + w * ncell].reshape(ncell, w).squeeze().copy()
        o += w * ncell
    return out


def read_jac_out(path, niter, nnode):
    a = np.frombuffer(open(path, "rb").read()[16:], np.float64)
    return {"u_sum": a[:niter].copy(), "u_max": a[niter:2 * niter].copy(),
            "u": a[2 * niter:2 * niter + nnode].copy(),
            "du": a[2 * niter + nnode:2 * niter + 2 * nnode].copy()}


def plan_fixture(tmp, mesh, prob, real_bytes, part):
    mb, jb, ob = (os.path.join(tmp, n) for n in ("m.bin", "j.bin", "plans.bin"))
    M.write_mesh_bin(mesh, mb)
    M.write_jacobi_bin(prob, jb)
    args = [os.path.join(REFBIN, "plan_dump"), mb, jb, str(real_bytes), ob]
    if part:
        args.append(f"OP_PART_SIZE={part}")
    run(args)
    raw = open(ob, "rb").read()
    # loc_maps follow each plan's ind_maps; re-parse with known arg layouts
    layouts = [[1, 1, 1, 1, 0, 0], [1] * 8, [1, 1, 1, 1, 1, 0], [0, 1, 1, 0]]
    sizes = [mesh.ncell, mesh.nedge, mesh.nbedge, prob.nedge]
    out = {}
    o = 0
    for k, (name, lay, n) in enumerate(zip(["adt_calc", "res_calc", "bres_calc", "jac_res"],
                                           layouts, sizes)):
        hdr = np.frombuffer(raw[o:o + 24], np.int32)
        nb, nc, nsh, nargs, ninds, nn = (int(v) for v in hdr)
        o += 24
        tr = np.frombuffer(raw[o:o + 8], np.float32)
        o += 8
        pre = f"{name}_"

        def take(cnt, dt, key):
            nonlocal o
            dt = np.dtype(dt)
            out[pre + key] = np.frombuffer(raw[o:o + cnt * dt.itemsize], dt).copy()
            o += cnt * dt.itemsize
            return out[pre + key]

        out[pre + "header"] = np.array([nb, nc, nsh, nargs, ninds, nn], np.int32)
        out[pre + "transfer"] = tr.copy()
        take(nb, np.int32, "nthrcol")
        take(nn, np.int32, "thrcol")
        take(nb, np.int32, "offset")
        take(nb, np.int32, "nelems")
        take(nc, np.int32, "ncolblk")
        take(nb, np.int32, "blkmap")
        take(nb * ninds, np.int32, "ind_offs")
        take(nb * ninds, np.int32, "ind_sizes")
        nind = take(ninds, np.int32, "nindirect")
        for m in range(ninds):
            take(int(nind[m]), np.int32, f"ind_map{m}")
        for a in range(nargs):
            if lay[a]:
                take(nn, np.int16, f"loc_map{a}")
    assert o == len(raw), "plan dump not fully parsed"
    return out


def airfoil_small(tmp):
    """24x12 random O-mesh, 20 iterations, every form and schedule."""
    mesh = M.o_mesh(24, 12, "random", seed=1)
    mb, cb, ob = (os.path.join(tmp, n) for n in ("m.bin", "c.bin", "o.bin"))
    M.write_mesh_bin(mesh, mb)
    niter = 20
    fx = {}
    for form, consts in (("f64", M.flow_constants()), ("f32", M.flow_constants_f32_reference())):
        c = consts.as_array()
        c.astype("<f8").tofile(cb)
        fx[f"{form}_consts"] = c
        run([os.path.join(REFBIN, f"airfoil_seq_{form}"), mb, cb, str(niter), ob])
        for k, v in read_airfoil_out(ob, niter, mesh.ncell).items():
            fx[f"{form}_seq_{k}"] = v
        run([os.path.join(REFBIN, f"airfoil_omp_{form}"), mb, cb, str(niter), ob,
             "OP_PART_SIZE=64"], threads=4)
        for k, v in read_airfoil_out(ob, niter, mesh.ncell).items():
            fx[f"{form}_omp_p64_t4_{k}"] = v
    fx["niter"] = np.array(niter)
    fx["mesh"] = np.array([24, 12, 1])
    np.savez_compressed(os.path.join(GOLD, "airfoil_24x12_random_s1.npz"), **fx)


def airfoil_180k(tmp):
    """600x300 random O-mesh, FP64, 1000 iterations, sequential order."""
    mesh = M.o_mesh(600, 300, "random", seed=1)
    mb, cb, ob = (os.path.join(tmp, n) for n in ("m.bin", "c.bin", "o.bin"))
    M.write_mesh_bin(mesh, mb)
    c = M.flow_constants().as_array()
    c.astype("<f8").tofile(cb)
    niter = 1000
    run([os.path.join(REFBIN, "airfoil_seq_f64"), mb, cb, str(niter), ob])
    o = read_airfoil_out(ob, niter, mesh.ncell)
    rng = np.random.Generator(np.random.PCG64(7))
    sample = np.sort(rng.choice(mesh.ncell, 512, replace=False))
    np.savez_compressed(
        os.path.join(GOLD, "airfoil_600x300_random_s1_f64.npz"),
        rms=o["rms"], consts=c, sample=sample, q_sample=o["q"][sample],
        adt_sample=o["adt"][sample], q_sum=o["q"].sum(axis=0),
        q_sumsq=(o["q"] ** 2).sum(axis=0), niter=np.array(niter),
        mesh=np.array([600, 300, 1]))


def airfoil_unchanged(tmp):
    """The reference airfoil program as shipped (float, new_grid.dat, rms
    printed every 100 iterations) on a 120x60 Cartesian O-mesh."""
    mesh = M.o_mesh(120, 60, "cartesian")
    M.write_grid_dat(mesh, os.path.join(tmp, "new_grid.dat"))
    out = run([os.path.join(REFBIN, "airfoil_ref")], cwd=tmp)
    with open(os.path.join(GOLD, "airfoil_ref_120x60_cartesian.txt"), "w") as f:
        f.write(out)


def jacobi(tmp):
    prob = M.jacobi_grid(64, seed=1)
    jb, ob = os.path.join(tmp, "j.bin"), os.path.join(tmp, "o.bin")
    M.write_jacobi_bin(prob, jb)
    niter = 200
    fx = {"pp": prob.pp, "A": prob.A, "r": prob.r}
    run([os.path.join(REFBIN, "jac_seq_f64"), jb, "0.25", "-0.25", str(niter), ob])
    for k, v in read_jac_out(ob, niter, prob.nnode).items():
        fx[f"seq_{k}"] = v
    run([os.path.join(REFBIN, "jac_omp_f64"), jb, "0.25", "-0.25", str(niter), ob,
         "OP_PART_SIZE=1024"], threads=4)
    for k, v in read_jac_out(ob, niter, prob.nnode).items():
        fx[f"omp_p1024_t4_{k}"] = v
    fx["niter"] = np.array(niter)
    np.savez_compressed(os.path.join(GOLD, "jacobi_64_random_s1_f64.npz"), **fx)
    out = run([os.path.join(REFBIN, "jac_ref")])
    with open(os.path.join(GOLD, "jac1_ref_nn6.txt"), "w") as f:
        f.write(out)
    # jac2: the same problem with multi-component, mixed-type dats
    out = run([os.path.join(REFBIN, "jac2_ref")])
    with open(os.path.join(GOLD, "jac2_ref_nn6.txt"), "w") as f:
        f.write(out)


def plans(tmp):
    mesh = M.o_mesh(24, 12, "random", seed=1)
    prob = M.jacobi_grid(16, seed=1)
    fx = {}
    for rb in (4, 8):
        for part in (0, 64):
            for k, v in plan_fixture(tmp, mesh, prob, rb, part).items():
                fx[f"r{rb}_p{part}_{k}"] = v
    np.savez_compressed(os.path.join(GOLD, "plans_24x12_random_s1.npz"), **fx)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    a = ap.parse_args()
    build_ref()
    with tempfile.TemporaryDirectory() as tmp:
        plans(tmp)
        airfoil_small(tmp)
        jacobi(tmp)
        airfoil_unchanged(tmp)
        if not a.quick:
            airfoil_180k(tmp)
    print("fixtures written to", GOLD)


if __name__ == "__main__":
    main()
