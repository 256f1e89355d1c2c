"""SI exports of full steady solves against the reference's published G0001 samples.

The solver executable runs G0001 on LF1 and LF2 to steady (FP64, default configuration) through
the command-line driver, which writes the SI exports.  They are compared with the reference's
published converged-state samples (docs/data/G0001_{LF1,LF2}.npz, fixture G0001_fields.npz):
per control volume p, temperature, Mach and speed on LF1, per wall fragment x, patch and Cp on
both levels, and the published CL / CD / CM_z.
"""
import json

import numpy as np
import pytest

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import cli

from .common import GOLDEN, pins

pytestmark = pytest.mark.gpu


def test_steady_exports_match_published_samples(tmp_path):
    rc = cli.main(["run", "--points", str(GOLDEN / "airfoils" / "G0001_airfoil_points.csv"), "--levels", "LF1", "LF2",
                   "--out", str(tmp_path)])
    ref = np.load(GOLDEN / "G0001_fields.npz")
    published = {r["profile"]: r for r in pins()["full_steady_G0001"]}
    assert rc == 0
    for level in ("LF1", "LF2"):
        out = tmp_path / level / "output"
        meta = json.loads((out / "metadata.json").read_text())
        print(level, meta["SI_EOS_relative_error"], meta["total_cells"], flush=True)
        assert meta["total_cells"] == int(ref[f"{level}_cells"]) and meta["surface_load_integration_passed"]
        surface = np.load(out / "surface.npz")
        assert np.array_equal(surface["patch_id"], ref[f"{level}_patch"])
        assert np.allclose(surface["x"], ref[f"{level}_surface_x"], rtol=0, atol=1e-12)
        cp_err = float(np.max(np.abs(surface["Cp"] - ref[f"{level}_cp"])))
        print(level, "max |dCp|", cp_err, flush=True)
        assert cp_err <= 1e-8
        loads = json.loads((out / "loads.json").read_text())["total"]
        for key, pub_key in (("CL", "cl"), ("CD", "cd"), ("CM_z", "cm")):
            assert abs(loads[key] - published[level][pub_key]) <= 1e-8, key
    field = np.load(tmp_path / "LF1" / "output" / "field.npz")
    for key in ("p", "temperature", "mach"):
        err = float(np.max(np.abs(field[key] - ref[f"LF1_{key}"]) / np.abs(ref[f"LF1_{key}"])))
        print("LF1", key, "max relative error", err, flush=True)
        assert err <= 1e-8, key
    speed = np.hypot(field["u"], field["v"])
    assert float(np.max(np.abs(speed - ref["LF1_speed"]))) <= 1e-8 * float(np.max(ref["LF1_speed"]))
