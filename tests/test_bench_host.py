"""bench.py's host-side helpers (no GPU): the lookup of measured traffic in
the committed counter summary and the algorithmic bytes of the stage
kernels, on the 2400x1200 workload the summary was measured on."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from oplus2_matlab_amd import op2  # noqa: E402


def test_measured_traffic_reads_the_rows_of_the_tiles_configuration():
    stage = {"strategy": 5, "part_size": 256, "block_size": 256}
    res = {"strategy": 5, "part_size": 512, "block_size": 512}
    si = bench.measured_traffic(2400, 1200, stage, ["stage/7si"])
    sq = bench.measured_traffic(2400, 1200, stage, ["stage/5sq"])
    assert si == pytest.approx(512.53e6) and sq == pytest.approx(420.08e6)
    assert bench.measured_traffic(2400, 1200, stage, ["stage/7si", "stage/5sq"]) == \
        pytest.approx((si + sq) / 2)
    assert bench.measured_traffic(2400, 1200, res, ["res_calc"]) == pytest.approx(386.60e6)
    # unprofiled configurations, meshes and keys give None
    assert bench.measured_traffic(2400, 1200, {"strategy": 5, "part_size": 128,
                                               "block_size": 128}, ["res_calc"]) is None
    assert bench.measured_traffic(600, 300, stage, ["stage/7si"]) is None
    assert bench.measured_traffic(2400, 1200, stage, ["stage/9"]) is None
    assert bench.measured_traffic(2400, 1200, None, ["res_calc"]) is None


def test_stage_bytes_of_the_profiled_kinds():
    """The algorithmic bytes per launch quoted in the traffic summary."""
    from oplus2_matlab_amd import mesh as M
    mesh = M.o_mesh(2400, 1200, "cartesian", seed=1)
    si = op2.KIND_BRES | op2.KIND_SAVE | op2.KIND_ADT | op2.KIND_ADT_SIDES | \
        op2.KIND_SAVE_IN_PLACE
    sq = op2.KIND_BRES | op2.KIND_ADT | op2.KIND_ADT_SIDES | op2.KIND_QOLD_IS_Q
    assert bench.stage_bytes(mesh, si, 8) == pytest.approx(466.6368e6, rel=1e-6)
    assert bench.stage_bytes(mesh, sq, 8) == pytest.approx(374.4768e6, rel=1e-6)


def test_stage_tiles_option(monkeypatch):
    monkeypatch.setattr(sys, "argv", ["bench.py", "--stage-tiles=-1/0", "--steps", "3"])
    a = bench.parse()
    assert a.stage_tiles == "-1/0" and a.steps == 3
    monkeypatch.setattr(sys, "argv", ["bench.py"])
    assert bench.parse().stage_tiles == "0/0"
