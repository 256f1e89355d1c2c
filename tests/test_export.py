"""SI exports (export.py) on the CPU: closed forms, load bookkeeping and the certification checks.

Artefacts are written the way the solver executable writes them (cells 13 columns, walls 11,
wall_stress 14, viscous_loads.json), from a real LF1 mesh of the host mesher and a uniform
freestream state, whose SI values are known in closed form: p = p_inf, T = T_inf,
Mach = M_inf, rho = p_inf / (R T_inf), |V| = M_inf a_inf.
"""
import json
import math

import numpy as np
import pytest

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import config as CF
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import export as EX

from .common import case_mesh, example_polygon


@pytest.fixture(scope="module")
def lf1():
    case, mesh, params = case_mesh("G0001", "LF1")
    arrays = mesh.arrays()
    c = arrays["cells"]
    w = arrays["walls"]
    cells = np.column_stack([c["ix"], c["iy"], c["span"], c["level"], c["x0"], c["x1"], c["y0"], c["y1"],
                             c["volume"], c["cx"], c["cy"], c["wall_begin"], c["wall_count"]]).astype(np.float64)
    walls = np.column_stack([w["owner_cell"], w["source_edge_id"], w["body_id"], w["patch_id"], w["x0"], w["y0"],
                             w["x1"], w["y1"], w["nx"], w["ny"], w["length"]]).astype(np.float64)
    state = np.asarray(mesh.initialize(params, case.freestream), dtype=np.float64).reshape(-1, len(cells))
    return case, cells, walls, state


def write_run(tmp_path, lf1, viscous_scale=1e-3, seed=0, corrupt=False):
    """Artefacts of a run at the freestream state: owner pressure on every fragment, a seeded
    small viscous force, loads summed as the executable sums them (moments about the origin)."""
    case, cells, walls, state = lf1
    raw = tmp_path / "flow"
    np.save(str(raw) + ".cells.npy", cells)
    np.save(str(raw) + ".walls.npy", walls)
    np.save(str(raw) + ".final.state.npy", state)
    rng = np.random.default_rng(seed)
    nw = len(walls)
    p = np.full(nw, 1.0 / case.gamma)  # the freestream pressure
    ln, nx, ny = walls[:, 10], walls[:, 8], walls[:, 9]
    x, y = 0.5 * (walls[:, 4] + walls[:, 6]), 0.5 * (walls[:, 5] + walls[:, 7])
    fxp, fyp = p * nx * ln, p * ny * ln
    fxv, fyv = viscous_scale * rng.standard_normal(nw) * ln, viscous_scale * rng.standard_normal(nw) * ln
    stress = np.column_stack([np.arange(nw), walls[:, 0], walls[:, 0], np.full(nw, 0.01), np.full(nw, 2e-7),
                              -fxv / ln, -fyv / ln, (nx * -fyv - ny * -fxv) / ln, p, fxp, fyp, fxv, fyv,
                              x * fyv - y * fxv])
    np.save(str(raw) + ".final.wall_stress.npy", stress)
    rows = np.column_stack([fxp, fyp, x * fyp - y * fxp, fxv, fyv, x * fyv - y * fxv])
    total = rows.sum(axis=0)
    if corrupt:
        total[0] += 1e-3
    patches = [{"patch": int(pid), "loads": rows[walls[:, 3] == pid].sum(axis=0).tolist()}
               for pid in np.unique(walls[:, 3])]
    (tmp_path / "flow.final.viscous_loads.json").write_text(json.dumps({"total": total.tolist(), "patches": patches}))
    return raw, rows


def test_freestream_fields_match_the_closed_form(tmp_path, lf1):
    raw, _ = write_run(tmp_path, lf1)
    c = CF.merged()
    f = c["freestream"]
    field, cells = EX.field_data(raw, c)
    sc = CF.scales(c)
    assert np.allclose(field["p"], f["p_Pa"], rtol=1e-14, atol=0)
    assert np.allclose(field["temperature"], f["T_K"], rtol=1e-14, atol=0)
    assert np.allclose(field["mach"], f["Mach"], rtol=1e-14, atol=0)
    assert np.allclose(field["rho"], f["p_Pa"] / (f["R_J_kgK"] * f["T_K"]), rtol=1e-14, atol=0)
    speed = np.hypot(field["u"], field["v"])
    assert np.allclose(speed, f["Mach"] * math.sqrt(f["gamma"] * f["R_J_kgK"] * f["T_K"]), rtol=1e-14, atol=0)
    assert np.allclose(np.arctan2(field["v"], field["u"]), math.radians(f["AOA_degrees"]), rtol=0, atol=1e-14)
    assert np.array_equal(field["x"], cells[:, 9] * sc["length"])
    # The scales themselves (reference docs: a = 295.068 m/s, q = 35740.4355 Pa, Re_c = 1.033e7).
    assert abs(sc["velocity"] - 295.068) < 5e-4 and abs(sc["q_inf"] - 35740.4355) < 1e-3
    assert abs(sc["Re_chord"] / 1.033e7 - 1) < 1e-3


def test_loads_rotation_moment_transfer_and_surface_sum(tmp_path, lf1):
    raw, rows = write_run(tmp_path, lf1)
    c = CF.merged()
    loads = EX.body_loads(raw, c)
    sc = CF.scales(c)
    force = sc["pressure"] * sc["length"]
    fx, fy, mz = rows[:, 0].sum() + rows[:, 3].sum(), rows[:, 1].sum() + rows[:, 4].sum(), rows[:, 2].sum() + rows[:, 5].sum()
    t = loads["total"]
    assert math.isclose(t["force_x"], fx * force, rel_tol=1e-12, abs_tol=1e-9)
    # Lift and drag are the force rotated by the angle of attack: the magnitude is kept.
    assert math.isclose(math.hypot(t["lift"], t["drag"]), math.hypot(t["force_x"], t["force_y"]), rel_tol=1e-12)
    alpha = math.radians(c["freestream"]["AOA_degrees"])
    assert math.isclose(t["drag"], math.cos(alpha) * t["force_x"] + math.sin(alpha) * t["force_y"], rel_tol=1e-12,
                        abs_tol=1e-9)
    # Moment about (0.25, 0) from the moment about the origin: M_ref = M_o - x_r F_y + y_r F_x.
    assert math.isclose(t["moment_z"], (mz - 0.25 * fy) * force * sc["length"], rel_tol=1e-12, abs_tol=1e-9)
    assert math.isclose(t["CL"], t["lift"] / (sc["q_inf"] * sc["length"]), rel_tol=1e-12, abs_tol=1e-15)
    # Parts add up, per patch too.
    for key in EX.LOAD_KEYS:
        assert math.isclose(t[key], loads["pressure"][key] + loads["viscous"][key], rel_tol=1e-12, abs_tol=1e-9)
    assert set(loads["patches"]) == {str(int(p)) for p in np.unique(lf1[2][:, 3])}
    # A uniform pressure on the closed body gives no net pressure force.
    assert abs(loads["pressure"]["force_x"]) < 1e-9 * force and abs(loads["pressure"]["force_y"]) < 1e-9 * force


def test_export_case_writes_certified_products(tmp_path, lf1):
    raw, _ = write_run(tmp_path, lf1)
    c = CF.merged({"output": {"sample_points_xy_c": [[1.5, 0.3], [-2.0, 1.0]], "region_xy_c": [-1.0, 3.0, -1.0, 1.0]}})
    polygon, _ = example_polygon("G0001")
    meta = EX.export_case(raw, tmp_path / "output", c, polygon)
    out = tmp_path / "output"
    assert meta["positivity_passed"] and meta["surface_load_integration_passed"]
    assert meta["SI_EOS_relative_error"] <= 2e-12
    field = np.load(out / "field.npz")
    assert list(field.files) == ["cell_id", "fluid_area_m2", "x", "y", "p", "rho", "u", "v", "mach", "temperature"]
    assert meta["selected_cells"] == len(field["cell_id"]) < meta["total_cells"]
    header = (out / "field.csv").read_text().splitlines()[0]
    assert header == ",".join(field.files)
    surface = np.load(out / "surface.npz")
    assert np.allclose(surface["Cp"], 0.0, atol=1e-12)  # freestream pressure on the wall
    assert set(np.unique(surface["patch_id"])) <= {1.0, 2.0, 3.0}
    probes = np.load(out / "probes.npz")
    cells = lf1[1]
    for i, (px, py) in enumerate([(1.5, 0.3), (-2.0, 1.0)]):
        k = int(np.argmin(np.hypot(cells[:, 9] - px, cells[:, 10] - py)))
        assert probes["nearest_cell"][i] == k
    aero = np.load(out / "aerodynamics.npz")
    assert list(aero.files) == CF.DEFAULT_LOAD_FIELDS
    loads_csv = (out / "loads.csv").read_text().splitlines()
    assert loads_csv[0].startswith("component,force_x_N_per_m") and [r.split(",")[0] for r in loads_csv[1:]] == [
        "pressure", "viscous", "total"]
    assert json.loads((out / "loads.json").read_text())["q_inf_Pa"] == CF.scales(c)["q_inf"]
    assert (out / "mesh.svg").exists() is False  # no agglomeration table was written


def test_certification_rejects_mismatched_loads_and_bad_probes(tmp_path, lf1):
    raw, _ = write_run(tmp_path, lf1, corrupt=True)
    polygon, _ = example_polygon("G0001")
    with pytest.raises(EX.ExportError, match="integration mismatch"):
        EX.export_case(raw, tmp_path / "output", CF.merged(), polygon)
    raw, _ = write_run(tmp_path, lf1)
    with pytest.raises(EX.ExportError, match="inside/on the airfoil"):
        EX.export_case(raw, tmp_path / "o2", CF.merged({"output": {"sample_points_xy_c": [[0.5, 0.0]]}}), polygon)
    with pytest.raises(EX.ExportError, match="outside the open computational domain"):
        EX.export_case(raw, tmp_path / "o3", CF.merged({"output": {"sample_points_xy_c": [[20.0, 0.0]]}}), polygon)
    with pytest.raises(ValueError, match="output.fields"):
        EX.export_case(raw, tmp_path / "o4", CF.merged({"output": {"fields": ["x", "x"]}}), polygon)
    bad = np.load(str(raw) + ".final.state.npy")
    bad[0, 7] = -1.0
    np.save(str(raw) + ".final.state.npy", bad)
    with pytest.raises(EX.ExportError, match="nonphysical"):
        EX.export_case(raw, tmp_path / "o5", CF.merged(), polygon)
