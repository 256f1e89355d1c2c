"""The solver executable and its command-line driver (no GPU: mesh mode).

The settings file carries the reference pipeline's keys (pipeline/configuration.py:151-167) and
the executable writes the reference's artefacts: NPY v1 <f8 arrays with the reference column
counts (cells 13, faces 9, walls 11, state nv x n), mesh.json, snapshot JSON, result.json with
the geometric audit and gpu.json.
"""
import json

import numpy as np
import pytest

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import cli
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import config as CF

from .common import GOLDEN, pins

REFERENCE_KEYS = {"name", "mode", "backend", "mesh_preset", "viscosity", "turbulence", "sa_layout", "stop_time",
                  "budget", "mu_star", "mach", "aoa", "gamma", "events", "nx", "ny", "levels", "buffer", "max_cells",
                  "xmin", "xmax", "ymin", "ymax", "cfl", "fixed_dt", "flux", "moment_x", "moment_y", "raw_tol",
                  "load_tol", "hold", "tmin", "period", "res_xmin", "res_xmax", "res_ymin", "res_ymax",
                  "outlet_p_star", "sa_nu_inf", "sutherland_star", "viscous_cfl", "sa_cfl", "wall_probe_min_h",
                  "bc0", "bc1", "bc2", "bc3"}


def test_settings_carry_the_reference_keys():
    text = cli.settings_text(CF.merged(), "LF2", "G0001", "hip")
    rows = [line.split() for line in text.splitlines()]
    keys = {r[0] for r in rows if r[0] != "box"}
    assert REFERENCE_KEYS <= keys and keys - REFERENCE_KEYS == {"model", "precision"}
    boxes = [r for r in rows if r[0] == "box"]
    assert [[float(v) for v in b[1:5]] + [int(b[5])] for b in boxes] == CF.PRESETS["LF2"]["boxes"]
    values = {r[0]: r[1] for r in rows if r[0] != "box"}
    assert float(values["mach"]) == 1.65 and values["levels"] == "6" and values["backend"] == "hip"
    # mu* = mu_inf / (rho a c): the same nondimensionalisation as config.make_case
    assert float(values["mu_star"]) == CF.make_case("LF2").mu_star


@pytest.mark.parametrize("profile", ["LF1"])
def test_mesh_mode_writes_reference_artefacts(tmp_path, profile):
    rc = cli.main(["run", "--points", str(GOLDEN / "airfoils" / "G0001_airfoil_points.csv"), "--levels", profile,
                   "--mode", "mesh", "--out", str(tmp_path)])
    assert rc == 0
    prefix = tmp_path / profile / "flow"
    result = json.loads(prefix.with_suffix(".result.json").read_text())
    expected = {(r["case"], r["profile"]): r["cells"] for r in pins()["short_runs_t0.01"]}[("G0001", profile)]
    assert result["fluid_control_volumes"] == expected and result["stop_reason"] == "mesh-only"
    assert result["area_error"] < 1e-10 and result["wall_length_error"] < 1e-11 and result["closure_relative_linf"] < 1e-9
    cells = np.load(prefix.with_suffix(".cells.npy"))
    faces = np.load(prefix.with_suffix(".faces.npy"))
    walls = np.load(prefix.with_suffix(".walls.npy"))
    state = np.load(tmp_path / profile / "flow.initial.state.npy")
    assert cells.shape == (expected, 13) and faces.shape[1] == 9 and walls.shape[1] == 11
    assert state.shape == (5, expected) and state.dtype == np.float64
    assert np.allclose(state[0], 1.0) and np.allclose(state[3], 1 / 1.4 / 0.4 + 0.5 * 1.65 ** 2)
    mesh = json.loads(prefix.with_suffix(".mesh.json").read_text())
    assert mesh["state_shape"] == [5, expected] and mesh["levels"] == 5
    gpu = json.loads(prefix.with_suffix(".gpu.json").read_text())
    assert gpu["allocated_device_bytes"] == 0
    # Viscous snapshot of the uniform initial state: 14 wall-stress columns per fragment; the
    # uniform pressure integrates to zero force on the closed body.
    stress = np.load(tmp_path / profile / "flow.initial.wall_stress.npy")
    assert stress.shape == (walls.shape[0], 14) and np.array_equal(stress[:, 1], walls[:, 0])
    loads = json.loads((tmp_path / profile / "flow.initial.viscous_loads.json").read_text())
    assert abs(loads["total"][0]) < 1e-12 and abs(loads["total"][1]) < 1e-12
    assert {p["patch"] for p in loads["patches"]} == {1, 2, 3}
    sa = np.load(tmp_path / profile / "flow.initial.sa.npy")
    assert sa.shape == (expected, 3) and np.all(sa[:, 2] > 0)
