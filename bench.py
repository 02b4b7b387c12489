#!/usr/bin/env python3
"""Airfoil benchmark on one MI355X: cell-updates/s and res_calc HBM GB/s.

A step is one airfoil iteration (save_soln, then twice adt_calc, res_calc,
bres_calc, update with the rms reduction; airfoil/airfoil.cpp:197-257) over
the 2400 x 1200 O-mesh (2.88M cells, 5.76M edges), run through the OP2 C ABI
with every dat resident in HBM.  value = cells x steps / seconds of the
timed steps.

roofline: res_calc, the hot loop.  achieved = algorithmic bytes of one call
(each touched dat once per direction plus each map once: 76 B/edge FP64)
divided by the mean device time of one call, from HIP events recorded on the
runtime's stream around every call in the timed steps.

cpu_baseline: the reference's OpenMP schedule (the oracle's bit-exact
restatement of airfoil/<loop>_kernel.cpp on the reference plan) on this
host's cores, over a bounded number of iterations of the same mesh.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "airfoil cell-updates/s & res_calc HBM GB/s (FP64 vs FP32 form) on 1 MI355X"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mesh", default="2400x1200", help="NIxNJ O-mesh cells")
    ap.add_argument("--ordering", default="cartesian", choices=["cartesian", "random", "rcm"])
    ap.add_argument("--dtype", default="double", choices=["double", "float"])
    ap.add_argument("--strategy", default="atomic", choices=["atomic", "coloured"])
    ap.add_argument("--part", type=int, default=0, help="res_calc part size (0 = default)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the cpu_baseline sample (0 = skip)")
    return ap.parse_args()


def cpu_baseline(mesh, consts, target_s):
    import numpy as np

    from oracle import oracle

    nthr = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    # one iteration sizes the sample; loop_seconds excludes plan construction
    _, _, secs = oracle.airfoil_omp(mesh, consts, 1, "f64", nthreads=nthr)
    per_iter = max(float(np.sum(secs)), 1e-3)
    iters = int(max(2, min(200, target_s / per_iter)))
    _, _, secs = oracle.airfoil_omp(mesh, consts, iters, "f64", nthreads=nthr)
    t = float(np.sum(secs))
    return {"value": mesh.ncell * iters / t, "unit": "cell-updates/s", "cores": nthr,
            "kind": "port",
            "sample": f"{iters} iterations of the same FP64 airfoil mesh on the reference "
                      f"OpenMP schedule (oracle restatement, bit-exact to the reference, "
                      f"adaptive part size), {nthr} threads, {t:.1f} s",
            "res_calc_gbs": None}


def main():
    a = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); multi-GPU sharding of "
              "the mesh is a design note, not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from oplus2_matlab_amd import airfoil as AF
    from oplus2_matlab_amd import mesh as M
    from oplus2_matlab_amd import op2

    ni, nj = (int(v) for v in a.mesh.lower().split("x"))
    mesh = M.o_mesh(ni, nj, a.ordering, seed=1)
    consts = M.flow_constants().as_array()
    strategy = {"atomic": op2.STRATEGY_ATOMIC, "coloured": op2.STRATEGY_COLOURED}[a.strategy]

    op2.op_init([], 0)
    for loop in ("res_calc", "bres_calc"):
        op2.set_loop_config(loop, strategy, a.part if loop == "res_calc" else 0, 0)
    af = AF.Airfoil(mesh, consts, a.dtype)
    for _ in range(a.warmup):
        af.iterate()
    op2.sync()
    op2.timing_reset()
    t0 = time.perf_counter()
    hist = [af.iterate() for _ in range(a.steps)]
    op2.sync()
    dt = time.perf_counter() - t0

    loops = {}
    for name in ("save_soln", "adt_calc", "res_calc", "bres_calc", "update"):
        st = op2.loop_stats(name)
        if st:
            n, secs, bytes_ = st
            loops[name] = {"calls": n, "us_per_call": 1e6 * secs / n,
                           "GBs": bytes_ * n / secs / 1e9, "bytes_per_call": bytes_}
    rc = loops["res_calc"]
    achieved = rc["GBs"]
    sreal = 8 if a.dtype == "double" else 4
    result = {
        "metric": METRIC,
        "value": mesh.ncell * a.steps / dt,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * dt / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.dtype == "double" else "f32 state/flux, f64 update+rms",
        "data": "synthetic",
        "config": {"workload": f"airfoil {a.dtype} {ni}x{nj} O-mesh ({mesh.ncell} cells, "
                               f"{mesh.nedge} edges), {a.ordering} ordering",
                   "cells": mesh.ncell, "edges": mesh.nedge, "ordering": a.ordering,
                   "strategy": a.strategy, "part_size": a.part or None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "res_calc", "bytes_per_launch": rc["bytes_per_call"],
                     "bytes_per_edge": rc["bytes_per_call"] / mesh.nedge,
                     "us_per_launch": rc["us_per_call"]},
        "loops": loops,
        "rms_last": hist[-1],
    }
    op2.op_exit()
    if a.cpu_seconds > 0:
        result["cpu_baseline"] = cpu_baseline(mesh, consts, a.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    del np
    print(json.dumps(result))


if __name__ == "__main__":
    main()
