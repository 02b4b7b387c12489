#!/usr/bin/env python3
"""Airfoil benchmark on one MI355X: cell-updates/s and res_calc HBM GB/s.

A step is one airfoil iteration (save_soln, then twice adt_calc, res_calc,
bres_calc, update with the rms reduction; airfoil/airfoil.cpp:197-257) over
the 2400 x 1200 O-mesh (2.88M cells, 5.76M edges), with every dat resident
in HBM.  The iteration is captured once as a hipGraph with deferred
reductions (the rms of every iteration lands in the runtime's device log)
and replayed; value = cells x steps / seconds of the timed replays (FP64).
Each graph holds --graph-iters iterations (default 50: every graph boundary
costs a separate save_soln + adt_calc kernel and one stored res), and the capture
fuses loops that can share a pass: update with the next iteration's
save_soln and each update or save_soln with the adt_calc that follows
(--no-fusion captures the loops one kernel each).

roofline: the stage kernel, the dominant kernel of the timed graphs (with
stage fusion, one kernel per Runge-Kutta stage runs res_calc, bres_calc,
update and the next save_soln and adt_calc).  achieved = algorithmic bytes
of its launches (stage_bytes: each map and dat it touches once per
direction) divided by their device time, from event-record nodes that the
runtime places around every 7th stage kernel inside the timed graphs
(op2_set_graph_kernel_timing), read for the last replay of each graph.
res_calc_roofline: the same for res_calc alone (76 B/edge FP64), from HIP
events the runtime records on its stream around every loop call of an eager
timed region; each of those calls follows an update, so the runtime takes
res as zero and does not read it, and res_read_skipped gives the fraction
on the bytes the call needs (without that read).  traffic = measured HBM bytes per launch from the committed
rocprofv3 counter summary, when this run's configuration is the profiled
one.

fp32_form: the same mesh with float state and flux (update and rms in
double): its throughput, speed-up over FP64, and its error against the FP64
run after --accuracy-iters iterations (rms history relative error and
max |q - q64|).

mesh_600x300: the 180k-cell O-mesh with seed-random numbering (renumbered
per --renumber, res_calc per --strategy as for the main mesh), 200 graph
steps.

orderings: the same 2.88M-cell mesh with every set numbered by a seed-random
permutation, once renumbered by the runtime (op2_renumber) and once as
given; cell-updates/s and per-loop GB/s of each (50 graph steps).

batched: 64 meshes of 10k-400k cells with their own mach and alpha
(7.12M cells) advanced as one union mesh, each boundary edge reading its
mesh's free stream and update summing each mesh's rms; the iteration is a
hipGraph as above.  value in cell-updates/s and its ratio to the single mesh,
for FP64 and the FP32 form.

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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "airfoil cell-updates/s & res_calc HBM GB/s (FP64 vs FP32 form) on 1 MI355X"
HBM_PEAK_GBS = 8000.0
# every 7th stage kernel of a timed graph is timed by event-record nodes
# (each node costs a few microseconds of the replay; an odd interval samples
# both stages of an iteration)
MARK_EVERY = 7
STRATEGIES = {"auto": 0, "coloured": 1, "atomic": 2, "direct": 3, "owner": 4, "owner-gather": 5,
              "tune": -1, "tune-offline": -1}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mesh", default="2400x1200", help="NIxNJ O-mesh cells")
    ap.add_argument("--ordering", default="cartesian", choices=["cartesian", "random", "rcm"])
    ap.add_argument("--strategy", default="tune", choices=list(STRATEGIES),
                    help="res_calc strategy; 'tune' runs the runtime's online autotuner in "
                         "the warmup iterations, 'tune-offline' its offline one")
    ap.add_argument("--part", type=int, default=0, help="res_calc part size (0 = default)")
    ap.add_argument("--block", type=int, default=0,
                    help="res_calc threads per workgroup with a fixed --strategy (0 = default)")
    ap.add_argument("--renumber", choices=["on", "off"], default="on",
                    help="renumber the sets for locality inside the runtime (op2_renumber)")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-form measurement")
    ap.add_argument("--accuracy-iters", type=int, default=1000,
                    help="iterations of the FP32-vs-FP64 accuracy run (0 = skip)")
    ap.add_argument("--no-orderings", dest="orderings", action="store_false",
                    help="skip the random-ordering measurements")
    ap.add_argument("--no-small-mesh", dest="small_mesh", action="store_false",
                    help="skip the 600x300 (180k-cell) measurement")
    ap.add_argument("--batch-meshes", type=int, default=64,
                    help="meshes of the batched measurement (0 = skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the cpu_baseline sample (0 = skip)")
    ap.add_argument("--stage-tiles", default="0/0",
                    help="PART/BLOCK of the stage kernels' tiles (0/0 = the runtime's default, "
                         "256/256; --stage-tiles=-1/0 = res_calc's configuration)")
    ap.add_argument("--no-fusion", dest="fusion", action="store_false",
                    help="capture the graphs without loop fusion")
    ap.add_argument("--graph-iters", type=int, default=50,
                    help="iterations per captured graph (at most the step count; a second "
                         "graph runs the remainder)")
    ap.add_argument("--no-kernel-marks", dest="kernel_marks", action="store_false",
                    help="capture the timed graphs without the event-record nodes that time "
                         "their stage kernels")
    return ap.parse_args()


TRAFFIC_PROFILE = os.path.join("profiles", "airfoil_2400x1200_fp64_traffic.md")


def measured_traffic(ni, nj, cfg, keys):
    """Mean HBM bytes per launch over `keys` (rows of the committed rocprofv3
    counter summary: FETCH_SIZE x copy-kernel factor + WRITE_SIZE, in MB),
    when this run is a profiled configuration (2400x1200 FP64, cartesian
    ordering renumbered, and `cfg`, the strategy, part and block size of the
    tiles the keyed kernels ran on, on one of the summary's `config:`
    lines); else None."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if (ni, nj) != (2400, 1200) or not cfg or not os.path.exists(path):
        return None
    # keyed rows follow the `config:` line of the runs they come from
    rows, section = {}, None
    for line in open(path):
        if line.startswith("config:"):
            section = tuple(int(v) for v in line.split()[1:4])
        elif line.startswith("| key:"):
            f = [c.strip() for c in line.split("|")]
            rows[(section, f[1][4:].strip().strip("`"))] = float(f[-4]) * 1e6
    got = (cfg.get("strategy"), cfg.get("part_size"), cfg.get("block_size"))
    if not keys or any((got, k) not in rows for k in keys):
        return None
    return sum(rows[(got, k)] for k in keys) / len(keys)


def capture_graphs(af, graph_iters, nsteps):
    """Graphs that run nsteps iterations: one of k = min(graph_iters, nsteps)
    iterations (replayed nsteps // k times) and, when k does not divide
    nsteps, one of the remaining nsteps % k iterations, both captured before
    any timing.  Returns (k, graph id, remainder graph id or None)."""
    k = max(1, min(graph_iters, nsteps))
    gid = af.capture_iteration(k)
    rid = af.capture_iteration(nsteps % k) if nsteps % k else None
    return k, gid, rid


def warm_graphs(af, k, gid, rid, warmup):
    """Untimed replays of the graphs the timed region launches (at least
    `warmup` iterations, and each graph at least once: a graph's first launch
    costs more than the later ones)."""
    from oplus2_matlab_amd import op2

    op2.graph_launch(gid, max(1, -(-warmup // k)))
    if rid is not None:
        op2.graph_launch(rid, 1)
    op2.sync()
    op2.reduction_log()


def measure(mesh, consts, dtype, a, renumber=None, fixed=None, steps=None, marks=False):
    """Eager timed region (per-loop HIP events) then graph timed region.
    renumber: override of --renumber; fixed: (res_calc strategy, part) instead
    of --strategy; steps: override of --steps; marks: time the graphs' stage
    kernels with event-record nodes."""
    from oplus2_matlab_amd import airfoil as AF
    from oplus2_matlab_amd import op2

    ren = (a.renumber == "on") if renumber is None else renumber
    nsteps = a.steps if steps is None else steps
    op2.op_init([], 0)
    try:
        if fixed:
            op2.set_loop_config("res_calc", fixed[0], fixed[1], 0)
        elif STRATEGIES[a.strategy] >= 0 and (a.strategy != "auto" or a.part or a.block):
            op2.set_loop_config("res_calc", max(STRATEGIES[a.strategy], 0), a.part, a.block)
        af = AF.Airfoil(mesh, consts, dtype, renumber=ren)
        tuned = None
        if not fixed:
            tuned = (af.autotune(reps=3, online=True) if a.strategy == "tune" else
                     af.autotune() if a.strategy == "tune-offline" else None)
        for _ in range(a.warmup):
            af.iterate()
        op2.sync()
        op2.timing_reset()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            af.iterate()
        op2.sync()
        dt_eager = time.perf_counter() - t0
        loops = {}
        for name in ("save_soln", "adt_calc", "res_calc", "bres_calc", "update"):
            st = op2.loop_stats(name)
            if st:
                n, secs, bytes_ = st
                loops[name] = {"calls": n, "us_per_call": 1e6 * secs / n,
                               "GBs": bytes_ * n / secs / 1e9, "bytes_per_call": bytes_}
        op2.set_graph_kernel_timing(MARK_EVERY if marks else 0)
        k, gid, rid = capture_graphs(af, a.graph_iters, nsteps)
        op2.set_graph_kernel_timing(0)
        warm_graphs(af, k, gid, rid, a.warmup)
        t0 = time.perf_counter()
        op2.graph_launch(gid, nsteps // k)
        if rid is not None:
            op2.graph_launch(rid, 1)
        op2.sync()
        dt_graph = time.perf_counter() - t0
        # the stage kernels of the last replay of each graph (event-record
        # nodes inside the graphs, on the runtime stream)
        stages = op2.graph_kernel_times(gid) + (op2.graph_kernel_times(rid) if rid is not None
                                                else [])
        log = op2.reduction_log()
        rms_last = float((log[-1] / mesh.ncell) ** 0.5) if log.size else None
        op2.set_deferred_reductions(False)
    finally:
        op2.op_exit()
        op2.set_loop_config(None, 0, 0, 0)
    return {"value": mesh.ncell * nsteps / dt_graph, "ms_per_step": 1e3 * dt_graph / nsteps,
            "eager_ms_per_step": 1e3 * dt_eager / nsteps, "loops": loops, "rms_last": rms_last,
            "tuned": tuned, "graph_iters": k, "stages": stages}


def stage_bytes(mesh, kind, size):
    """Algorithmic bytes of one stage kernel of the given kind (op2 KIND_*
    bits) with `size`-byte reals: each map and dat it touches once per
    direction.  Always: pedge and pecell (8 B per edge each), x (2 reals per
    node), q, adt and qold read and the new q and adt written (4+1+4+4+1
    reals per cell), and for bres_calc pbedge, pbecell and bound (16 B per
    boundary edge); SAVE writes qold (4 reals per cell; none when the new q
    is written over qold in place of its own store; a stage whose qold is
    that q reads it once), ADT reads pcell (16
    B per cell) or, in the side form, the side recipes (2 B per cell), and
    res (4 reals per cell) is read and/or stored as the kind says.  The rms
    partials (8 B per tile) are left out."""
    from oplus2_matlab_amd import op2

    b = 16 * mesh.nedge + 2 * size * mesh.nnode + 14 * size * mesh.ncell
    if kind & op2.KIND_QOLD_IS_Q:  # qold and q are the same record: read once
        b -= 4 * size * mesh.ncell
    if kind & op2.KIND_BRES:
        b += 16 * mesh.nbedge
    if kind & op2.KIND_SAVE and not kind & op2.KIND_SAVE_IN_PLACE:
        b += 4 * size * mesh.ncell
    if kind & op2.KIND_ADT:
        b += (2 if kind & op2.KIND_ADT_SIDES else 16) * mesh.ncell
    if kind & op2.KIND_READS_RES:
        b += 4 * size * mesh.ncell
    if kind & op2.KIND_STORES_RES:
        b += 4 * size * mesh.ncell
    return b


def stage_roofline(mesh, stages, size):
    """Stage kernels of the timed graphs: algorithmic bytes and device time
    summed over the bracketed launches of each graph's last replay."""
    if not stages:
        return None
    nbytes = sum(stage_bytes(mesh, kind, size) for kind, _ in stages)
    secs = sum(ms for _, ms in stages) * 1e-3
    by_kind = {}
    for kind, ms in stages:
        by_kind.setdefault(kind, []).append(ms)
    return {"launches": len(stages), "bytes_per_launch": nbytes / len(stages),
            "us_per_launch": 1e6 * secs / len(stages), "GBs": nbytes / secs / 1e9,
            "by_kind": {str(kd): {"launches": len(v), "us_per_launch": 1e3 * sum(v) / len(v),
                                  "bytes_per_launch": stage_bytes(mesh, kd, size)}
                        for kd, v in sorted(by_kind.items())}}


def measure_batch(a, batch, dtype):
    """64 meshes (10k-400k cells, own mach and alpha) advanced together: one
    graph-captured iteration of the union mesh, replayed; per-mesh rms in the
    device log."""
    from oplus2_matlab_amd import airfoil as AF
    from oplus2_matlab_amd import op2

    op2.op_init([], 0)
    try:
        if STRATEGIES[a.strategy] >= 0 and (a.strategy != "auto" or a.part or a.block):
            op2.set_loop_config("res_calc", max(STRATEGIES[a.strategy], 0), a.part, a.block)
        af = AF.AirfoilBatch(batch, dtype, renumber=a.renumber == "on")
        tuned = (af.autotune(reps=3, online=True) if a.strategy == "tune" else
                 af.autotune() if a.strategy == "tune-offline" else None)
        for _ in range(a.warmup):
            af.iterate()
        op2.sync()
        op2.timing_reset()
        for _ in range(10):
            af.iterate()
        loops = {}
        for name in ("save_soln", "adt_calc", "res_calc", "bres_calc", "update"):
            st = op2.loop_stats(name)
            if st:
                loops[name] = {"calls": st[0], "us_per_call": 1e6 * st[1] / st[0]}
        k, gid, rid = capture_graphs(af, a.graph_iters, a.steps)
        warm_graphs(af, k, gid, rid, a.warmup)
        # the reduction log holds 2 x 64 results per iteration: replays are
        # read back every `per` graphs (a sync) so that it never overflows
        per = max(1, op2.REDUCTION_LOG_CAPACITY // (2 * batch.nmesh * k))
        t0 = time.perf_counter()
        done = 0
        while done < a.steps // k:
            n = min(per, a.steps // k - done)
            op2.graph_launch(gid, n)
            log = op2.reduction_log()
            done += n
        if rid is not None:
            op2.graph_launch(rid, 1)
            log = op2.reduction_log()
        op2.sync()
        dt = time.perf_counter() - t0
        last = np.sqrt(log.reshape(-1, 2, batch.nmesh)[-1, 1] / af.mesh_cells)
        op2.set_deferred_reductions(False)
    finally:
        op2.op_exit()
        op2.set_loop_config(None, 0, 0, 0)
    ncell = batch.union.ncell
    return {"value": ncell * a.steps / dt, "unit": "cell-updates/s",
            "ms_per_step": 1e3 * dt / a.steps, "meshes": batch.nmesh, "cells": ncell,
            "edges": batch.union.nedge,
            "cells_min": int(min(m.ncell for m in batch.meshes)),
            "cells_max": int(max(m.ncell for m in batch.meshes)),
            "rms_last_min": float(last.min()), "rms_last_max": float(last.max()),
            "graph_iters": k, "tuned": tuned, "loops": loops}


def batch_accuracy(batch, iters, renumber, graph_iters):
    """The batch in the FP32 form against FP64 after `iters` iterations
    (fused graphs, as timed): per-mesh rms-history relative error and
    max |q - q64| relative to the mesh's max |q|, worst over the meshes."""
    from oplus2_matlab_amd import airfoil as AF
    from oplus2_matlab_amd import op2

    out = {}
    for dtype in ("double", "float"):
        op2.op_init([], 0)
        try:
            out[dtype] = AF.run_airfoil_batch(batch, iters, dtype, graph=True, renumber=renumber,
                                              graph_iters=max(1, min(graph_iters, iters - 1)))
        finally:
            op2.op_exit()
    (h64, q64), (h32, q32) = out["double"], out["float"]
    worst = worst_q = 0.0
    for i in range(batch.nmesh):
        lo, hi = batch.cell_off[i], batch.cell_off[i + 1]
        worst = max(worst, float(np.max(np.abs(h32[:, i] - h64[:, i]) / np.abs(h64[:, i]))))
        worst_q = max(worst_q, float(np.max(np.abs(q32[lo:hi].astype(np.float64) - q64[lo:hi])) /
                                     np.max(np.abs(q64[lo:hi]))))
    return {"iterations": iters, "meshes": batch.nmesh,
            "rms_history_max_rel_err": worst, "max_abs_q_minus_q64_rel": worst_q}


def accuracy(mesh, consts, iters, renumber, graph_iters=1):
    """FP32 form against FP64 after `iters` iterations, replayed as graphs of
    `graph_iters` fused iterations (the path the throughput figures time)."""
    from oplus2_matlab_amd import airfoil as AF
    from oplus2_matlab_amd import op2

    out = {}
    for dtype in ("double", "float"):
        op2.op_init([], 0)
        try:
            out[dtype] = AF.run_airfoil(mesh, consts, iters, dtype, graph=True,
                                        renumber=renumber,
                                        graph_iters=max(1, min(graph_iters, iters - 1)))
        finally:
            op2.op_exit()
    h64, q64 = out["double"][0], out["double"][1]
    h32, q32 = out["float"][0], out["float"][1]
    return {"iterations": iters, "graph_iters": max(1, min(graph_iters, iters - 1)),
            "rms_history_max_rel_err": float(np.max(np.abs(h32 - h64) / np.abs(h64))),
            "rms_final_rel_err": float(abs(h32[-1] - h64[-1]) / abs(h64[-1])),
            "max_abs_q_minus_q64": float(np.max(np.abs(q32.astype(np.float64) - q64)))}


def host_cpus():
    """The host's CPU model (/proc/cpuinfo, as lscpu reports it), its logical
    CPU count, the CPUs this process may run on, and the cgroup CPU quota
    (cpu.max; None when unlimited or unreadable)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    return {"model": model, "logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)), "cgroup_cpu_quota": quota}


def cpu_baseline(mesh, consts, target_s):
    from oracle import oracle

    nthr = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    host = host_cpus()
    # five iterations size the sample (the first, cold one alone costs about
    # two later ones); loop_seconds excludes plan construction
    _, _, secs = oracle.airfoil_omp(mesh, consts, 5, "f64", nthreads=nthr)
    per_iter = max(float(np.sum(secs)) / 5, 1e-3)
    iters = int(max(2, min(2000, target_s / per_iter)))
    _, _, secs = oracle.airfoil_omp(mesh, consts, iters, "f64", nthreads=nthr)
    t = float(np.sum(secs))
    return {"value": mesh.ncell * iters / t, "unit": "cell-updates/s", "cores": nthr,
            "kind": "port", "host": host,
            "sample": f"{iters} iterations of the same FP64 airfoil mesh on the reference "
                      f"OpenMP schedule (oracle restatement, bit-exact to the reference, "
                      f"adaptive part size), {nthr} threads, {t:.1f} s",
            "res_calc_GBs": float(mesh.nedge * 76.0 * iters * 2 / secs[2] / 1e9)}


def main():
    a = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); multi-GPU sharding of "
              "the mesh is a design note, not implemented", file=sys.stderr)
        sys.exit(2)

    from oplus2_matlab_amd import mesh as M
    from oplus2_matlab_amd import op2

    op2.set_graph_fusion(a.fusion)
    st_part, st_block = (int(v) for v in a.stage_tiles.split("/"))
    op2.set_stage_config(st_part, st_block)
    ni, nj = (int(v) for v in a.mesh.lower().split("x"))
    mesh = M.o_mesh(ni, nj, a.ordering, seed=1)
    consts = M.flow_constants().as_array()

    r64 = measure(mesh, consts, "double", a, marks=a.kernel_marks)
    rc = r64["loops"]["res_calc"]
    res_cfg = (r64["tuned"] or {}).get("res_calc") if a.strategy.startswith("tune") else \
        {"strategy": STRATEGIES[a.strategy], "part_size": a.part, "block_size": a.block or 256}
    profiled = a.renumber == "on" and a.ordering == "cartesian"
    res_traffic = measured_traffic(ni, nj, res_cfg, ["res_calc"]) if profiled else None
    res_roof = {"bound": "hbm", "achieved": rc["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": rc["GBs"] / HBM_PEAK_GBS, "traffic": res_traffic,
                "traffic_source": TRAFFIC_PROFILE if res_traffic else None,
                "kernel": "res_calc (eager calls)", "bytes_per_launch": rc["bytes_per_call"],
                "bytes_per_edge": rc["bytes_per_call"] / mesh.nedge,
                "us_per_launch": rc["us_per_call"]}
    # every timed res_calc follows an update, which left res zero: the
    # runtime does not read it (the algorithmic bytes count that read)
    needed = rc["bytes_per_call"] - 4 * 8 * mesh.ncell
    res_roof["res_read_skipped"] = {
        "bytes_per_launch": needed,
        "frac": needed / (rc["us_per_call"] * 1e-6) / 1e9 / HBM_PEAK_GBS}
    stage = stage_roofline(mesh, r64["stages"], 8)
    stage_cfg = dict(res_cfg or {}) if st_part < 0 else \
        {"strategy": STRATEGIES["owner-gather"], "part_size": st_part or 256,
         "block_size": (st_block or 256) if st_part else 256}
    if stage:
        traffic = measured_traffic(ni, nj, stage_cfg,
                                   [f"stage/{kd & 7}" + ("s" if kd & op2.KIND_ADT_SIDES else "") +
                                    ("i" if kd & op2.KIND_SAVE_IN_PLACE else "") +
                                    ("q" if kd & op2.KIND_QOLD_IS_Q else "")
                                    for kd, _ in r64["stages"]]) \
            if profiled else None
        roofline = {"bound": "hbm", "achieved": stage["GBs"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": stage["GBs"] / HBM_PEAK_GBS, "traffic": traffic,
                    "traffic_source": TRAFFIC_PROFILE if traffic else None,
                    "kernel": "k_stage (one Runge-Kutta stage: res_calc, bres_calc, update, "
                              "save_soln, adt_calc), in the timed graphs",
                    "bytes_per_launch": stage["bytes_per_launch"],
                    "us_per_launch": stage["us_per_launch"], "launches": stage["launches"],
                    "by_kind": stage["by_kind"]}
    else:
        roofline = res_roof
    result = {
        "metric": METRIC,
        "value": r64["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"airfoil FP64 {ni}x{nj} O-mesh ({mesh.ncell} cells, "
                               f"{mesh.nedge} edges), {a.ordering} ordering, iteration "
                               f"replayed as a hipGraph",
                   "cells": mesh.ncell, "edges": mesh.nedge, "ordering": a.ordering,
                   "strategy": a.strategy, "part_size": a.part or None,
                   "renumbered": a.renumber == "on", "graph_fusion": a.fusion,
                   "graph_iters": r64["graph_iters"],
                   "stage_tiles": {k: stage_cfg.get(k) for k in ("part_size", "block_size")},
                   "tuned": r64["tuned"]},
        "roofline": roofline,
        "res_calc_roofline": res_roof,
        "eager_ms_per_step": r64["eager_ms_per_step"],
        "loops": r64["loops"],
        "rms_last": r64["rms_last"],
    }
    if not a.no_fp32:
        r32 = measure(mesh, consts, "float", a, marks=a.kernel_marks)
        result["fp32_form"] = {"value": r32["value"], "unit": "cell-updates/s",
                               "ms_per_step": r32["ms_per_step"],
                               "speedup_vs_fp64": r32["value"] / r64["value"],
                               "res_calc_GBs": r32["loops"]["res_calc"]["GBs"],
                               "res_calc_us": r32["loops"]["res_calc"]["us_per_call"],
                               "stage_kernels": stage_roofline(mesh, r32["stages"], 4),
                               "tuned": r32["tuned"],
                               "loops": r32["loops"]}
        if a.accuracy_iters > 0:
            result["fp32_form"]["accuracy"] = accuracy(mesh, consts, a.accuracy_iters,
                                                       a.renumber == "on", a.graph_iters)
    if a.small_mesh:
        # the 600 x 300 O-mesh (180k cells) with seed-random numbering
        ms = M.o_mesh(600, 300, "random", seed=1)
        r = measure(ms, consts, "double", a, renumber=a.renumber == "on", steps=200)
        result["mesh_600x300"] = {
            "value": r["value"], "unit": "cell-updates/s", "ms_per_step": r["ms_per_step"],
            "cells": ms.ncell, "edges": ms.nedge, "ordering": "random",
            "renumbered": a.renumber == "on", "strategy": a.strategy, "tuned": r["tuned"],
            "loops": {k: {"us_per_call": v["us_per_call"], "GBs": v["GBs"]}
                      for k, v in r["loops"].items()}}
    if a.orderings:
        # the same mesh with a seed-random numbering of every set: renumbered
        # by the runtime, and as given (fixed res_calc strategies: the
        # fastest of each in sweeps, to keep the run short)
        mr = M.o_mesh(ni, nj, "random", seed=1)
        orders = {}
        for label, ren, fixed in (("random_renumbered", True, (5, 256)),
                                  ("random_as_given", False, (4, 128))):
            r = measure(mr, consts, "double", a, renumber=ren, fixed=fixed, steps=50)
            orders[label] = {"value": r["value"], "unit": "cell-updates/s",
                             "ms_per_step": r["ms_per_step"], "res_calc_strategy": fixed,
                             "loops": {k: {"us_per_call": v["us_per_call"], "GBs": v["GBs"]}
                                       for k, v in r["loops"].items()}}
        result["orderings"] = orders
    if a.batch_meshes > 0:
        batch = M.mesh_batch(nmesh=a.batch_meshes, seed=1)
        rb = measure_batch(a, batch, "double")
        rb["ratio_vs_single_mesh"] = rb["value"] / r64["value"]
        if not a.no_fp32:
            rb32 = measure_batch(a, batch, "float")
            rb["fp32_form"] = {"value": rb32["value"], "ms_per_step": rb32["ms_per_step"],
                               "speedup_vs_fp64": rb32["value"] / rb["value"],
                               "tuned": rb32["tuned"], "loops": rb32["loops"]}
            if a.accuracy_iters > 0:
                rb["fp32_form"]["accuracy"] = batch_accuracy(
                    batch, min(a.accuracy_iters, 100), a.renumber == "on", a.graph_iters)
            if "fp32_form" in result:
                rb["fp32_form"]["ratio_vs_single_mesh"] = rb32["value"] / result["fp32_form"]["value"]
        result["batched"] = rb
    result["cpu_baseline"] = cpu_baseline(mesh, consts, a.cpu_seconds) if a.cpu_seconds > 0 else None
    print(json.dumps(result))


if __name__ == "__main__":
    main()
