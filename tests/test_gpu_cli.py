"""The solver executable on the GPU: a smoke run (one check at t* = 0.01) of G0001 on LF1
through the command-line driver; the step count is the published one (41) and the artefacts
are complete."""
import csv
import json

import numpy as np
import pytest

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import cli

from .common import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [64, 32])
def test_smoke_run_through_the_executable(tmp_path, precision):
    rc = cli.main(["run", "--points", str(GOLDEN / "airfoils" / "G0001_airfoil_points.csv"), "--levels", "LF1",
                   "--mode", "smoke", "--precision", str(precision), "--out", str(tmp_path)])
    log = (tmp_path / "LF1" / "solver.log").read_text()
    print(log[-2000:], flush=True)
    assert rc == 0
    prefix = tmp_path / "LF1" / "flow"
    result = json.loads(prefix.with_suffix(".result.json").read_text())
    assert result["stop_reason"] == "short-validation" and result["checks"] == 1 and not result["failed"]
    assert result["steps"] == 41 and result["physical_time"] == 0.01 and result["precision"] == precision
    rows = list(csv.DictReader(open(prefix.with_suffix(".history.csv"))))
    assert len(rows) == 1 and int(rows[0]["steps"]) == 41 and float(rows[0]["Cl"]) == result["Cl"]
    state = np.load(tmp_path / "LF1" / "flow.final.state.npy")
    assert state.shape == (5, result["fluid_control_volumes"]) and np.all(np.isfinite(state))
    gpu = json.loads(prefix.with_suffix(".gpu.json").read_text())
    assert gpu["allocated_device_bytes"] > 0 and gpu["graph_nodes_per_trial"] > 0
