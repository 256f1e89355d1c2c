"""SI exports of a finished solve: cell fields, wall surface, probes and body loads.

The counterpart of the reference's pipeline/export.py (field_data :21-34, body_loads :37-62,
surface_data :65-94, probes :97-119, mesh_svg :122-132, export_case :135-200).  It reads the
solver executable's artefacts for a run prefix P (`P.final.state.npy`, `P.cells.npy`,
`P.walls.npy`, `P.final.wall_stress.npy`, `P.final.viscous_loads.json`, `P.agglomeration.npy`)
and writes into an output folder `field`, `surface`, `probes` and `aerodynamics` tables
(`.npz` and / or `.csv`), `loads.json` / `loads.csv`, `mesh.svg` and `metadata.json`, with the
reference's column names, units and formats so that downstream readers need no change.

Two certifications guard every export, as in the reference: the SI fields satisfy p = rho R T
(relative error <= 2e-12), and the wall fragments' pressure + viscous forces and moment sum to
the certified body loads (rtol 2e-11, atol 2e-8).  Nondimensionalisation: density, speed of
sound and chord are the reference scales (config.scales).
"""
from __future__ import annotations

import json
import math
import time
from pathlib import Path

import numpy as np

from . import config as CF

UNITS = dict(x="m", y="m", p="Pa", rho="kg/m^3", u="m/s", v="m/s", mach="1", temperature="K")

# Columns of the artefacts (cells.npy: 13, walls.npy: 11, wall_stress.npy: 14 per row).
CELL_AREA, CELL_CX, CELL_CY = 8, 9, 10
WALL_OWNER, WALL_EDGE, WALL_PATCH = 0, 1, 3
WALL_X0, WALL_Y0, WALL_X1, WALL_Y1, WALL_NX, WALL_NY, WALL_LEN = 4, 5, 6, 7, 8, 9, 10
ST_DONOR, ST_DIST, ST_MU, ST_TAU_T, ST_P = 2, 3, 4, 7, 8
ST_FXP, ST_FYP, ST_FXV, ST_FYV = 9, 10, 11, 12

LOAD_KEYS = ("force_x", "force_y", "drag", "lift", "moment_z", "CD", "CL", "CM_z")


class ExportError(ValueError):
    """A run whose artefacts fail a certification check (no product files are written)."""


def write_table(stem: Path, data: dict, formats) -> None:
    """`stem.npz` (one array per column) and / or `stem.csv` (header row, %.17g values)."""
    stem = Path(stem)
    if "npz" in formats:
        np.savez(str(stem) + ".npz", **data)
    if "csv" in formats:
        columns = np.column_stack([np.asarray(v, dtype=np.float64) for v in data.values()])
        np.savetxt(str(stem) + ".csv", columns, delimiter=",", header=",".join(data), comments="", fmt="%.17g")


def field_data(raw: str | Path, c: dict) -> tuple[dict, np.ndarray]:
    """SI cell-average fields at the final state: x, y, p, rho, u, v, mach, temperature."""
    sc = CF.scales(c)
    state = np.load(str(raw) + ".final.state.npy")
    cells = np.load(str(raw) + ".cells.npy")
    rho = state[0]
    with np.errstate(divide="ignore", invalid="ignore"):
        vx, vy = state[1] / rho, state[2] / rho
        p = (c["freestream"]["gamma"] - 1.0) * (state[3] - 0.5 * rho * (vx * vx + vy * vy))
    if not np.all(np.isfinite(state)) or np.any(rho <= 0) or np.any(p <= 0):
        raise ExportError("nonphysical or nonfinite state: exports are not certified")
    sound = np.sqrt(c["freestream"]["gamma"] * p / rho)
    data = dict(x=cells[:, CELL_CX] * sc["length"], y=cells[:, CELL_CY] * sc["length"], p=p * sc["pressure"],
                rho=rho * sc["rho"], u=vx * sc["velocity"], v=vy * sc["velocity"], mach=np.hypot(vx, vy) / sound,
                temperature=p / rho * sc["temperature"])
    return data, cells


def _load_part(fx: float, fy: float, mz: float, c: dict, sc: dict) -> dict:
    """SI force, lift / drag, moment about the reference point and coefficients of one part
    (fx, fy, mz nondimensional, moment about the origin)."""
    alpha = math.radians(c["freestream"]["AOA_degrees"])
    ca, sa = math.cos(alpha), math.sin(alpha)
    xr, yr = c["solver"]["moment_reference_xy_c"]
    moment = mz - xr * fy + yr * fx
    drag, lift = ca * fx + sa * fy, -sa * fx + ca * fy
    force = sc["pressure"] * sc["length"]
    coefficient = sc["pressure"] / sc["q_inf"]
    return dict(force_x=fx * force, force_y=fy * force, drag=drag * force, lift=lift * force,
                moment_z=moment * force * sc["length"], CD=drag * coefficient, CL=lift * coefficient,
                CM_z=moment * coefficient)


def body_loads(raw: str | Path, c: dict) -> dict:
    """Pressure, viscous and total body loads in SI, in total and per wall patch."""
    sc = CF.scales(c)
    loads = json.loads(Path(str(raw) + ".final.viscous_loads.json").read_text())

    def split(row):
        row = [float(v) for v in row]
        pressure, viscous = row[:3], row[3:6]
        return dict(pressure=_load_part(*pressure, c, sc), viscous=_load_part(*viscous, c, sc),
                    total=_load_part(*(a + b for a, b in zip(pressure, viscous)), c, sc))

    out = split(loads["total"])
    out["patches"] = {str(entry["patch"]): split(entry["loads"]) for entry in loads["patches"]}
    out["units"] = dict(force_x="N/m span", force_y="N/m span", drag="N/m span", lift="N/m span",
                        moment_z="N m/m span", CD="1", CL="1", CM_z="1")
    out["coordinates"] = ("x-y plane; positive Mz is counterclockwise; drag along freestream, lift rotated "
                          "+90 degrees")
    xr, yr = c["solver"]["moment_reference_xy_c"]
    out["moment_reference_m"] = [xr * sc["length"], yr * sc["length"]]
    out["integration"] = ("all body wall fragments; pressure and viscous parts; independent of field ROI and "
                          "selected surface patches; absolute static pressure")
    out["q_inf_Pa"] = sc["q_inf"]
    return out


def _wall_forces(walls: np.ndarray, stress: np.ndarray, c: dict, sc: dict) -> tuple:
    """Per-fragment SI force (pressure + viscous) and moment about the reference point."""
    scale = sc["pressure"] * sc["length"]
    fx = (stress[:, ST_FXP] + stress[:, ST_FXV]) * scale
    fy = (stress[:, ST_FYP] + stress[:, ST_FYV]) * scale
    x = 0.5 * (walls[:, WALL_X0] + walls[:, WALL_X1]) * sc["length"]
    y = 0.5 * (walls[:, WALL_Y0] + walls[:, WALL_Y1]) * sc["length"]
    xr, yr = np.asarray(c["solver"]["moment_reference_xy_c"], dtype=np.float64) * sc["length"]
    return fx, fy, (x - xr) * fy - (y - yr) * fx, x, y


def surface_data(raw: str | Path, c: dict, field: dict) -> dict:
    """Per wall fragment of the selected patches: ids, geometry, owner-cell state, Cp, donor
    viscosity, shear and forces in SI."""
    sc = CF.scales(c)
    walls = np.load(str(raw) + ".walls.npy")
    stress = np.load(str(raw) + ".final.wall_stress.npy")
    owner = walls[:, WALL_OWNER].astype(np.int64)
    nw = len(walls)
    fx, fy, moment, x, y = _wall_forces(walls, stress, c, sc)
    data = dict(wall_id=np.arange(nw), patch_id=walls[:, WALL_PATCH], source_edge_id=walls[:, WALL_EDGE],
                owner_cell=owner, donor_cell=stress[:, ST_DONOR], nx=walls[:, WALL_NX], ny=walls[:, WALL_NY],
                length_m=walls[:, WALL_LEN] * sc["length"])
    # Field columns on the wall: owner-cell p, rho and temperature; zero velocity and Mach on the
    # stationary no-slip wall.
    wall_fields = dict(x=("x", x), y=("y", y), p=("p_owner_Pa", field["p"][owner]),
                       rho=("rho_owner_kg_m3", field["rho"][owner]), u=("u_wall_m_s", np.zeros(nw)),
                       v=("v_wall_m_s", np.zeros(nw)), mach=("mach_wall", np.zeros(nw)),
                       temperature=("temperature_owner_K", field["temperature"][owner]))
    for key in c["output"]["fields"]:
        name, values = wall_fields[key]
        data[name] = values
    data["Cp"] = (stress[:, ST_P] * sc["pressure"] - c["freestream"]["p_Pa"]) / sc["q_inf"]
    data["mu_donor_Pa_s"] = stress[:, ST_MU] * sc["viscosity"]
    data["normal_distance_m"] = stress[:, ST_DIST] * sc["length"]
    data["wall_shear_signed_Pa"] = stress[:, ST_TAU_T] * sc["pressure"]
    scale = sc["pressure"] * sc["length"]
    data["pressure_force_x_N_per_m"] = stress[:, ST_FXP] * scale
    data["pressure_force_y_N_per_m"] = stress[:, ST_FYP] * scale
    data["viscous_force_x_N_per_m"] = stress[:, ST_FXV] * scale
    data["viscous_force_y_N_per_m"] = stress[:, ST_FYV] * scale
    data["moment_z_N_m_per_m"] = moment
    keep = np.isin(walls[:, WALL_PATCH], c["output"]["surface_patches"])
    return {key: np.asarray(value)[keep] for key, value in data.items()}


def _inside_or_on(polygon: np.ndarray, px: float, py: float) -> bool:
    """Even-odd test, plus distance <= 1e-12 to any edge counts as on the body."""
    a = np.asarray(polygon, dtype=np.float64)
    b = np.roll(a, 1, axis=0)
    d = b - a
    t = np.clip(((px - a[:, 0]) * d[:, 0] + (py - a[:, 1]) * d[:, 1]) / np.einsum("ij,ij->i", d, d), 0.0, 1.0)
    if np.any(np.hypot(px - a[:, 0] - t * d[:, 0], py - a[:, 1] - t * d[:, 1]) < 1e-12):
        return True
    crosses = (a[:, 1] > py) != (b[:, 1] > py)
    with np.errstate(divide="ignore", invalid="ignore"):
        xs = (b[:, 0] - a[:, 0]) * (py - a[:, 1]) / (b[:, 1] - a[:, 1]) + a[:, 0]
    return bool(np.count_nonzero(crosses & (px < xs)) % 2)


def probes(c: dict, field: dict, cells: np.ndarray, polygon: np.ndarray) -> dict | None:
    """Nearest fluid control volume (centroid, no interpolation) of each requested point."""
    points = c["output"]["sample_points_xy_c"]
    if not points:
        return None
    bx0, bx1, by0, by1 = c["mesh"]["bounds_xy_c"]
    chosen, distances = [], []
    for px, py in points:
        if not (bx0 < px < bx1 and by0 < py < by1):
            raise ExportError("probe point is outside the open computational domain")
        if _inside_or_on(polygon, px, py):
            raise ExportError("probe point lies inside/on the airfoil; use surface output for wall data")
        distance = np.hypot(cells[:, CELL_CX] - px, cells[:, CELL_CY] - py)
        k = int(np.argmin(distance))
        chosen.append(k)
        distances.append(distance[k])
    sc = CF.scales(c)
    pts = np.asarray(points, dtype=np.float64)
    return dict(query_x_m=pts[:, 0] * sc["length"], query_y_m=pts[:, 1] * sc["length"],
                nearest_cell=np.asarray(chosen), distance_m=np.asarray(distances) * sc["length"],
                **{key: field[key][chosen] for key in c["output"]["fields"]})


def mesh_svg(raw: str | Path, polygon: np.ndarray, path: Path) -> bool:
    """Agglomeration member rectangles around the airfoil and the body outline (layout sketch;
    written when the run has an agglomeration table)."""
    members_path = Path(str(raw) + ".agglomeration.npy")
    if not members_path.exists():
        return False
    members = np.load(members_path)
    x0, x1, y0, y1 = -0.6, 2.5, -0.65, 0.65
    shown = members[~((members[:, 3] < x0) | (members[:, 2] > x1) | (members[:, 5] < y0) | (members[:, 4] > y1))]
    rects = "".join(f'<rect x="{r[2]}" y="{-r[5]}" width="{r[3] - r[2]}" height="{r[5] - r[4]}"/>' for r in shown)
    outline = " ".join(f"{x},{-y}" for x, y in polygon)
    path.write_text(f'<svg xmlns="http://www.w3.org/2000/svg" viewBox="{x0} {-y1} {x1 - x0} {y1 - y0}">'
                    f'<rect x="{x0}" y="{-y1}" width="{x1 - x0}" height="{y1 - y0}" fill="white"/>'
                    f'<g fill="none" stroke="#6d9eb0" stroke-width=".0009">{rects}</g>'
                    f'<polygon points="{outline}" fill="white" stroke="#173745" stroke-width=".002"/></svg>')
    return True


def check_output_options(o: dict) -> None:
    """The output section's admissible values (reference pipeline/configuration.py:111-125)."""
    fields, load_fields = list(o["fields"]), list(o["load_fields"])
    if not fields or not set(fields) <= set(UNITS) or len(fields) != len(set(fields)):
        raise ValueError("output.fields must contain distinct supported field names")
    if not load_fields or not set(load_fields) <= set(CF.DEFAULT_LOAD_FIELDS) | {"moment"}:
        raise ValueError("unsupported output.load_fields")
    if not o["formats"] or not set(o["formats"]) <= {"csv", "npz"}:
        raise ValueError("output.formats must contain csv and/or npz")
    if not set(o["surface_patches"]) <= {1, 2, 3}:
        raise ValueError("surface_patches must use 1 upper, 2 lower, 3 trailing edge")
    if not all(isinstance(o[key], bool) for key in ("volume", "surface", "loads")):
        raise ValueError("output.volume/surface/loads must be booleans")
    region = o["region_xy_c"]
    if region is not None and not (len(region) == 4 and region[0] < region[1] and region[2] < region[3]):
        raise ValueError("output.region_xy_c must be [xmin, xmax, ymin, ymax]")
    for point in o["sample_points_xy_c"]:
        if len(point) != 2 or not all(math.isfinite(x) for x in point):
            raise ValueError("sample point requires two finite coordinates")


def export_case(raw: str | Path, out: str | Path, c: dict, polygon: np.ndarray) -> dict:
    """Certify the run's artefacts and write every product file into `out`; returns the
    metadata (also written as metadata.json)."""
    start = time.perf_counter()
    out = Path(out)
    o = c["output"]
    check_output_options(o)
    sc = CF.scales(c)
    field, cells = field_data(raw, c)
    keep = np.ones(len(cells), dtype=bool)
    if o["region_xy_c"] is not None:
        rx0, rx1, ry0, ry1 = o["region_xy_c"]
        keep = ((cells[:, CELL_CX] >= rx0) & (cells[:, CELL_CX] <= rx1) & (cells[:, CELL_CY] >= ry0)
                & (cells[:, CELL_CY] <= ry1))
    selected = dict(cell_id=np.flatnonzero(keep), fluid_area_m2=cells[keep, CELL_AREA] * sc["length"] ** 2,
                    **{key: field[key][keep] for key in o["fields"]})
    sampled = probes(c, field, cells, polygon)
    surface = surface_data(raw, c, field) if o["surface"] else None
    loads = body_loads(raw, c)
    # Certification over all fragments, whatever patches the surface output selects.
    walls = np.load(str(raw) + ".walls.npy")
    stress = np.load(str(raw) + ".final.wall_stress.npy")
    fx, fy, moment, _, _ = _wall_forces(walls, stress, c, sc)
    summed = [float(fx.sum()), float(fy.sum()), float(moment.sum())]
    certified = [loads["total"][key] for key in ("force_x", "force_y", "moment_z")]
    if not np.allclose(summed, certified, rtol=2e-11, atol=2e-8):
        raise ExportError("surface-to-load integration mismatch")
    eos = float(np.max(np.abs(field["p"] - field["rho"] * c["freestream"]["R_J_kgK"] * field["temperature"])
                       / field["p"]))
    if eos > 2e-12:
        raise ExportError("SI equation-of-state mismatch")
    out.mkdir(parents=True, exist_ok=True)
    for stem in ("field", "surface", "probes", "loads", "aerodynamics"):  # regenerate product files only
        for suffix in ("csv", "npz", "json"):
            (out / f"{stem}.{suffix}").unlink(missing_ok=True)
    if o["volume"]:
        write_table(out / "field", selected, o["formats"])
    if surface is not None:
        write_table(out / "surface", surface, o["formats"])
    if sampled is not None:
        write_table(out / "probes", sampled, o["formats"])
    if o["loads"]:
        (out / "loads.json").write_text(json.dumps(loads, indent=2) + "\n")
        total = loads["total"]
        flat = dict(fx=total["force_x"], fy=total["force_y"], drag=total["drag"], lift=total["lift"],
                    moment_z=total["moment_z"], moment=total["moment_z"], CD=total["CD"], CL=total["CL"],
                    CM_z=total["CM_z"])
        for part in ("pressure", "viscous"):
            for key in ("force_x", "force_y", "drag", "lift", "moment_z"):
                flat[f"{part}_{key}"] = loads[part][key]
        write_table(out / "aerodynamics", {key: np.asarray([flat[key]]) for key in o["load_fields"]}, o["formats"])
        if "csv" in o["formats"]:
            header = ("component,force_x_N_per_m,force_y_N_per_m,drag_N_per_m,lift_N_per_m,moment_z_N_m_per_m,"
                      "CD,CL,CM_z\n")
            rows = "".join(part + "," + ",".join(format(loads[part][key], ".17g") for key in LOAD_KEYS) + "\n"
                           for part in ("pressure", "viscous", "total"))
            (out / "loads.csv").write_text(header + rows)
    mesh_svg(raw, polygon, out / "mesh.svg")
    metadata = dict(
        fields=list(o["fields"]), volume_output=bool(o["volume"]), load_fields=list(o["load_fields"]),
        load_units="forces N/m span; moments N m/m span; CL/CD/CM dimensionless",
        units={key: UNITS[key] for key in o["fields"]}, selected_cells=int(keep.sum()), total_cells=len(cells),
        region_xy_c=o["region_xy_c"], selection="fluid control-volume centroid inside inclusive region; cell-average value",
        surface_sampling=("p/rho/temperature from wall owner; u=v=Mach=0 at stationary no-slip wall; "
                          "temperature_owner is not a prescribed wall temperature"),
        probes="nearest fluid-control-volume centroid, no interpolation; requested points inside/on body are rejected",
        patch_markers={"1": "upper", "2": "lower", "3": "vertical trailing edge"},
        load_scope="all airfoil fragments; independent of ROI/selected surface patches",
        positivity_passed=True, SI_EOS_relative_error=eos, surface_load_integration_passed=True,
        seconds=time.perf_counter() - start)
    (out / "metadata.json").write_text(json.dumps(metadata, indent=2) + "\n")
    return metadata
