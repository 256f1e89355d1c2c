"""Command-line driver of an airfoil solve: points -> settings -> solver executable -> results.

The counterpart of the reference's `quick_cfd.py run` (quick_cfd.py:219-231, settings from
pipeline/configuration.py:143-167): it prepares the polygon and patch files, writes the same
`settings.txt` keys (plus `model` and `precision`), runs the solver executable of the chosen
backend once per mesh level, reads back `result.json` / `gpu.json` and writes the SI exports
(export.py) of every viscous level into LEVEL/output.

  python -m quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.cli run \\
      --points G0001_airfoil_points.csv --levels LF1 LF2 --backend hip --out runs/G0001

Backends: `hip` runs this package's `quick_airfoil_hip` (the MI355X drop-in); `cpu` runs the
reference's own CPU build of `quick_airfoil`, whose path must be given with --cpu-executable
(this package ships no CPU solver).
"""
from __future__ import annotations

import argparse
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

from . import config as CF
from . import export as EX
from . import geometry as G

PACKAGE_DIR = Path(__file__).resolve().parent
HIP_EXECUTABLE = PACKAGE_DIR / "quick_airfoil_hip"

# Settings keys in the order of the reference pipeline (pipeline/configuration.py:151-167).
SETTING_KEYS = ["name", "mode", "backend", "mesh_preset", "viscosity", "turbulence", "sa_layout", "stop_time", "budget",
                "mu_star", "mach", "aoa", "gamma", "events", "nx", "ny", "levels", "buffer", "max_cells", "xmin", "xmax",
                "ymin", "ymax", "cfl", "fixed_dt", "flux", "moment_x", "moment_y", "raw_tol", "load_tol", "hold", "tmin",
                "period", "res_xmin", "res_xmax", "res_ymin", "res_ymax", "outlet_p_star", "sa_nu_inf",
                "sutherland_star", "viscous_cfl", "sa_cfl", "wall_probe_min_h", "bc0", "bc1", "bc2", "bc3"]


def settings_text(c: dict, preset: str, name: str, backend: str, mode: str | None = None, precision: int = 64) -> str:
    """settings.txt of one level: `key value` lines, then one `box x0 x1 y0 y1 level` per box."""
    f, p, s, m = c["freestream"], c["physics"], c["solver"], c["mesh"]
    sc = CF.scales(c)
    profile = CF.PRESETS[preset]
    cv = s["convergence"]
    mu = sc["mu_inf"] / sc["viscosity"]
    b, r = m["bounds_xy_c"], cv["residual_region_xy_c"]
    fixed = s["time_step"]["dt_seconds"] / sc["time"] if s["time_step"]["mode"] == "fixed" else 0
    values = dict(name=name, mode=mode or s["mode"], backend=backend, mesh_preset=preset, viscosity=p["viscosity"],
                  turbulence=p["turbulence"], sa_layout="fused", stop_time=s["stop_time_star"], budget=s["budget_seconds"],
                  mu_star=mu, mach=f["Mach"], aoa=f["AOA_degrees"], gamma=f["gamma"], events=1,
                  nx=m["base_shape"][0], ny=m["base_shape"][1], levels=profile["levels"], buffer=m["wall_buffer"],
                  max_cells=m["max_cells"], xmin=b[0], xmax=b[1], ymin=b[2], ymax=b[3], cfl=s["CFL"], fixed_dt=fixed,
                  flux=s["flux"], moment_x=s["moment_reference_xy_c"][0], moment_y=s["moment_reference_xy_c"][1],
                  raw_tol=cv["raw_tolerance"], load_tol=cv["coefficient_change"], hold=cv["hold"],
                  tmin=cv["min_time_star"], period=cv["check_period_star"], res_xmin=r[0], res_xmax=r[1],
                  res_ymin=r[2], res_ymax=r[3], outlet_p_star=c["boundaries"]["outlet_pressure_Pa"] / sc["pressure"],
                  sa_nu_inf=p["sa_nu_inf_over_nu"] * mu, sutherland_star=p["sutherland"]["S_K"] / sc["temperature"],
                  viscous_cfl=s["viscous_CFL"], sa_cfl=s["SA_CFL"], wall_probe_min_h=p["wall_probe_min_h"])
    for index, side in enumerate(["xmin", "xmax", "ymin", "ymax"]):
        values[f"bc{index}"] = c["boundaries"][side]
    lines = [f"{k} {float(values[k])!r}" if isinstance(values[k], float) else f"{k} {values[k]}" for k in SETTING_KEYS]
    lines += [f"model {p['model']}", f"precision {precision}"]
    boxes = [f"box {' '.join(repr(float(v)) if i < 4 else str(int(v)) for i, v in enumerate(bx))}" for bx in profile["boxes"]]
    return "\n".join(lines + boxes) + "\n"


def prepare_geometry(points_csv: Path, folder: Path) -> tuple[Path, Path, str]:
    """Polygon (airfoil.csv) and patch (patches.csv) files in the reference's order and format."""
    poly, patches = G.reorder_points(G.read_points(points_csv))
    folder.mkdir(parents=True, exist_ok=True)
    geometry = folder / "airfoil.csv"
    geometry.write_text("x,y\n" + "".join(f"{float(x)!r},{float(y)!r}\n" for x, y in poly))
    patch_file = folder / "patches.csv"
    patch_file.write_text("edge,patch\n" + "".join(f"{i},{int(p)}\n" for i, p in enumerate(patches)))
    stem = points_csv.stem
    name = stem.split("_")[0] if stem.upper().startswith("G") else stem
    return geometry, patch_file, name


def executable_for(backend: str, cpu_executable: str | None) -> Path:
    if backend == "hip":
        if not HIP_EXECUTABLE.exists():
            raise SystemExit(f"{HIP_EXECUTABLE} is not built; run __graft_entry__.build() first")
        return HIP_EXECUTABLE
    if not cpu_executable:
        raise SystemExit("--backend cpu runs the reference's CPU build of quick_airfoil: pass --cpu-executable PATH")
    return Path(cpu_executable)


def run(args) -> dict:
    overrides: dict = {}
    if args.config:
        overrides = json.loads(Path(args.config).read_text())
    cli = {"freestream": {}, "solver": {}, "physics": {}}
    for attr, section, key in (("mach", "freestream", "Mach"), ("aoa", "freestream", "AOA_degrees"),
                               ("cfl", "solver", "CFL"), ("stop_time", "solver", "stop_time_star"),
                               ("budget", "solver", "budget_seconds"), ("turbulence", "physics", "turbulence"),
                               ("viscosity", "physics", "viscosity"), ("model", "physics", "model"),
                               ("flux", "solver", "flux")):
        value = getattr(args, attr)
        if value is not None:
            cli[section][key] = value
    c = CF.merged({k: v for k, v in cli.items() if v}, CF.merged(overrides))
    out = Path(args.out)
    geometry, patch_file, name = prepare_geometry(Path(args.points), out / "geometry")
    exe = executable_for(args.backend, args.cpu_executable)
    summary = {"case": name, "backend": args.backend, "levels": {}}
    for preset in args.levels:
        folder = out / preset
        folder.mkdir(parents=True, exist_ok=True)
        settings = folder / "settings.txt"
        settings.write_text(settings_text(c, preset, name, args.backend, args.mode, args.precision))
        prefix = folder / "flow"
        cmd = [str(exe), "--settings", str(settings), "--geometry", str(geometry), "--patches", str(patch_file),
               "--out", str(prefix)]
        print("[run]", " ".join(cmd), file=sys.stderr, flush=True)
        done = subprocess.run(cmd, capture_output=True, text=True)
        (folder / "solver.log").write_text(done.stdout + done.stderr)
        result = json.loads(Path(f"{prefix}.result.json").read_text()) if Path(f"{prefix}.result.json").exists() else {}
        gpu = json.loads(Path(f"{prefix}.gpu.json").read_text()) if Path(f"{prefix}.gpu.json").exists() else {}
        summary["levels"][preset] = dict(returncode=done.returncode, result=result, gpu=gpu)
        # SI exports (reference quick_cfd.py:110): viscous runs that wrote a final state.
        if done.returncode == 0 and Path(f"{prefix}.final.viscous_loads.json").exists():
            polygon = np.loadtxt(geometry, delimiter=",", skiprows=1)
            summary["levels"][preset]["export"] = EX.export_case(prefix, folder / "output", c, polygon)
        print(f"[run] {preset}: rc={done.returncode} {result.get('stop_reason')} steps={result.get('steps')} "
              f"Cl={result.get('Cl')} Cd={result.get('Cd')} Cm={result.get('Cm')}", file=sys.stderr, flush=True)
    (out / "summary.json").write_text(json.dumps(summary, indent=2) + "\n")
    return summary


def main(argv=None):
    ap = argparse.ArgumentParser(description="airfoil points -> cut-cell mesh -> solver -> results")
    sub = ap.add_subparsers(dest="command", required=True)
    p = sub.add_parser("run", help="solve one airfoil on one or more mesh levels")
    p.add_argument("--points", required=True)
    p.add_argument("--config")
    p.add_argument("--out", required=True)
    p.add_argument("--levels", nargs="+", choices=sorted(CF.PRESETS), default=["LF1", "LF2"])
    p.add_argument("--backend", choices=["hip", "cpu"], default="hip")
    p.add_argument("--cpu-executable", help="the reference's CPU quick_airfoil (for --backend cpu)")
    p.add_argument("--mode", choices=["steady", "time", "smoke", "mesh", "rest"])
    p.add_argument("--precision", type=int, choices=[64, 32], default=64)
    for flag in ("mach", "aoa", "cfl", "stop_time", "budget"):
        p.add_argument("--" + flag.replace("_", "-"), dest=flag, type=float)
    p.add_argument("--turbulence", choices=["none", "sa-neg"])
    p.add_argument("--viscosity", choices=["constant", "sutherland"])
    p.add_argument("--model", choices=["laminar", "euler"])
    p.add_argument("--flux", choices=["hllc", "ausm"])
    args = ap.parse_args(argv)
    summary = run(args)
    return 0 if all(v["returncode"] == 0 for v in summary["levels"].values()) else 1


if __name__ == "__main__":
    sys.exit(main())
