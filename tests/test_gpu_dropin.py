"""The reference's own programs, unchanged, linked against this runtime.

oplus2_matlab_amd/lib/dropin/ holds airfoil/airfoil.cpp, airfoil_op.cpp,
jac1/jac.cpp and jac_op.cpp from the reference tree compiled against
include/ and linked with libop2hip.so (the csrc Makefile's `dropin` target,
run by build() where the reference sources are available).  Their printed
results are compared with what the reference's own build printed
(tests/golden/).  The float programs run this runtime's FP32 form (float
state and flux, double update and rms), so airfoil's rms may differ from the
all-float reference in the last printed digits.
"""

import os
import re
import subprocess

import pytest

from oplus2_matlab_amd import mesh as M

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
DROPIN = os.path.join(ROOT, "oplus2_matlab_amd", "lib", "dropin")


def need(name):
    """The built program; its absence is a failure (build() makes them from
    the reference sources, and the tree that reaches the GPU carries them)."""
    path = os.path.join(DROPIN, name)
    assert os.path.exists(path), f"{path} not built: run build() where the reference sources are"
    return path


def run(path, cwd, env=None):
    print(f"[gpu] running {os.path.basename(path)} {env or ''}", flush=True)
    full = dict(os.environ, **env) if env else None
    out = subprocess.run([path], cwd=cwd, capture_output=True, text=True, timeout=180, env=full)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return out.stdout


def rms_lines(text):
    vals = {}
    for ln in text.splitlines():
        m = re.match(r"^\s*(\d+)\s+([0-9.eE+-]+)\s*$", ln)
        if m:
            vals[int(m.group(1))] = float(m.group(2))
    return vals


def timing_table(text):
    """Rows of op_timing_output's table (common/op_lib_core.c:381-403):
    name -> (count, seconds, GB/s, GB/s or None)."""
    rows = {}
    lines = text.splitlines()
    start = next(i for i, ln in enumerate(lines) if "kernel name" in ln) + 2
    for ln in lines[start:]:
        f = ln.split()
        if len(f) == 5:
            rows[f[4]] = (int(f[0]), float(f[1]), float(f[2]), float(f[3]))
        elif len(f) == 4:
            rows[f[3]] = (int(f[0]), float(f[1]), float(f[2]), None)
        else:
            break
    return rows


@pytest.mark.parametrize("prog", ["airfoil", "airfoil_op"])
def test_unchanged_airfoil_program(tmp_path, golden_dir, prog):
    exe = need(prog)
    M.write_grid_dat(M.o_mesh(120, 60, "cartesian"), os.path.join(tmp_path, "new_grid.dat"))
    out = run(exe, tmp_path)
    got = rms_lines(out)
    ref = rms_lines(open(os.path.join(golden_dir, "airfoil_ref_120x60_cartesian.txt")).read())
    assert sorted(got) == sorted(ref) == list(range(100, 1001, 100))
    for it in ref:
        assert abs(got[it] - ref[it]) <= 2e-4 * abs(ref[it]), (it, got[it], ref[it])
    # the program's closing op_timing_output(): one row per loop with the
    # reference's columns; the loops through maps carry the second GB/s
    table = timing_table(out)
    print(f"[gpu] timing table {table}", flush=True)
    counts = {"save_soln": 1000, "adt_calc": 2000, "res_calc": 2000, "bres_calc": 2000,
              "update": 2000}
    assert {k: v[0] for k, v in table.items()} == counts
    for name, (_, secs, gbs, gbs2) in table.items():
        assert secs > 0 and gbs > 0, name
        assert (gbs2 is not None) == (name in ("adt_calc", "res_calc", "bres_calc")), name
        if gbs2 is not None:
            assert gbs2 >= gbs, name


@pytest.mark.parametrize("env", [None, "0"])
@pytest.mark.parametrize("prog", ["airfoil", "airfoil_op"])
def test_unchanged_airfoil_on_a_randomly_numbered_mesh(tmp_path, golden_dir, prog, env):
    """The same mesh written in a random numbering: renumbered by the
    runtime at the program's first res_calc (no call in the program), or
    run as given with OP2_RENUMBER=0; both give the golden rms history of
    the cartesian one."""
    exe = need(prog)
    M.write_grid_dat(M.o_mesh(120, 60, "random", seed=3), os.path.join(tmp_path, "new_grid.dat"))
    out = run(exe, tmp_path, env={"OP2_RENUMBER": env} if env else None)
    assert ("renumbering on map pecell (automatic)" in out) == (env is None)
    got = rms_lines(out)
    ref = rms_lines(open(os.path.join(golden_dir, "airfoil_ref_120x60_cartesian.txt")).read())
    assert sorted(got) == sorted(ref) == list(range(100, 1001, 100))
    for it in ref:
        assert abs(got[it] - ref[it]) <= 2e-4 * abs(ref[it]), (it, got[it], ref[it])


def test_unchanged_airfoil_random_numbering_runs_near_renumbered_speed(tmp_path):
    """The unchanged program on a 600 x 300 mesh written in a random
    numbering: with the default automatic renumbering its res_calc takes
    within 1.5x of the device time on the cartesian mesh; as given
    (OP2_RENUMBER=0) it is slower.  Times from the program's own
    op_timing_output table."""
    exe = need("airfoil")
    times = {}
    for label, ordering, env in (("cartesian", "cartesian", None), ("random", "random", None),
                                 ("random_as_given", "random", {"OP2_RENUMBER": "0"})):
        d = tmp_path / label
        d.mkdir()
        M.write_grid_dat(M.o_mesh(600, 300, ordering, seed=7), os.path.join(d, "new_grid.dat"))
        table = timing_table(run(exe, d, env=env))
        times[label] = table["res_calc"][1] / table["res_calc"][0]
    print(f"[gpu] res_calc seconds per call {times}", flush=True)
    assert times["random"] < 1.5 * times["cartesian"]
    assert times["random_as_given"] > times["random"]


@pytest.mark.parametrize("prog", ["jac", "jac_op"])
def test_unchanged_jac1_program(tmp_path, golden_dir, prog):
    exe = need(prog)
    out = run(exe, tmp_path)
    ref = open(os.path.join(golden_dir, "jac1_ref_nn6.txt")).read()
    want = [ln for ln in ref.splitlines() if "u max/rms" in ln]
    assert want and all(ln in out for ln in want)
    table = ref[ref.index("Results after"):].splitlines()[2:7]
    for ln in table:
        assert ln in out


@pytest.mark.parametrize("prog", ["jac2", "jac2_op"])
def test_unchanged_jac2_program_on_the_generic_path(tmp_path, golden_dir, prog):
    """jac2 (multi-component, mixed float/double dats) has no built-in
    kernel: its loops run on the generic path, compiled from the unchanged
    jac2/res.h and jac2/update.h (csrc/dropin/jac2_kernels.hip)."""
    exe = need(prog)
    out = run(exe, tmp_path)
    ref = open(os.path.join(golden_dir, "jac2_ref_nn6.txt")).read()
    want = [ln for ln in ref.splitlines() if "u max/rms" in ln]
    assert len(want) == 2 and all(ln in out for ln in want)
    table = ref[ref.index("Results after"):].splitlines()[2:7]
    for ln in table:
        assert ln in out
