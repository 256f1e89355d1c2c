"""Run configuration: physical defaults, mesh presets and nondimensionalisation.

Values follow the reference defaults (config/default.json) and its scaling
(pipeline/configuration.py:129-167): density, speed of sound and chord are the reference
scales, so the freestream is (rho, u, v, p) = (1, M cos a, M sin a, 1/gamma), temperature is
scaled by gamma*T_inf (T* = p*/rho* with R* = 1) and viscosity by rho*a*c.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field

# Reference config/default.json (freestream, physics, solver, mesh sections).
DEFAULTS = {
    "freestream": {"Mach": 1.65, "AOA_degrees": 1.5, "p_Pa": 18754.0, "T_K": 216.65, "chord_m": 1.0,
                   "gamma": 1.4, "R_J_kgK": 287.05},
    "boundaries": {"xmin": "farfield", "xmax": "farfield", "ymin": "farfield", "ymax": "farfield",
                   "outlet_pressure_Pa": 18754.0},
    "physics": {"model": "laminar", "turbulence": "sa-neg", "viscosity": "sutherland", "constant_mu_Pa_s": None,
                "sutherland": {"mu0_Pa_s": 1.716e-05, "T0_K": 273.15, "S_K": 110.4},
                "sa_nu_inf_over_nu": 3.0, "wall_probe_min_h": 0.25, "viscous_stabilization": "volume"},
    "solver": {"backend": "hip", "flux": "hllc", "CFL": 0.15, "viscous_CFL": 0.1, "SA_CFL": 0.1,
               "time_step": {"mode": "cfl", "dt_seconds": None}, "mode": "steady", "stop_time_star": 100.0,
               "budget_seconds": 118.0, "moment_reference_xy_c": [0.25, 0.0], "precision": 64,
               "convergence": {"raw_tolerance": 0.001, "coefficient_change": 0.005, "hold": 5,
                               "min_time_star": 3.0, "check_period_star": 0.5,
                               "residual_region_xy_c": [-1.0, 3.0, -1.0, 1.0]}},
    "mesh": {"bounds_xy_c": [-7.0, 14.0, -10.0, 10.0], "base_shape": [42, 40], "wall_buffer": 4,
             "max_cells": 2000000, "agglomerate_small": True},
    "output": {"fields": ["x", "y", "p", "rho", "u", "v", "mach", "temperature"], "formats": ["npz", "csv"],
               "region_xy_c": None, "sample_points_xy_c": [], "surface": True, "surface_patches": [1, 2, 3],
               "loads": True, "volume": True, "load_fields": None},
}

# Exported load columns (reference pipeline/configuration.py:9-10).
DEFAULT_LOAD_FIELDS = ["drag", "lift", "fx", "fy", "pressure_force_x", "pressure_force_y", "pressure_drag",
                       "pressure_lift", "viscous_force_x", "viscous_force_y", "viscous_drag", "viscous_lift",
                       "moment_z", "pressure_moment_z", "viscous_moment_z", "CD", "CL", "CM_z"]
DEFAULTS["output"]["load_fields"] = list(DEFAULT_LOAD_FIELDS)

# Mesh presets.  LF1/LF2 are the reference's (config/default.json "profiles"); L7/L8 refine the
# wall by one/two more levels with the LF2 boxes raised by the same amount.
PRESETS = {
    "LF1": {"levels": 5, "boxes": [[-0.5, 2.0, -0.75, 0.75, 3], [1.0, 4.0, -0.35, 0.65, 3], [-2.0, 6.0, -3.0, 3.0, 1]]},
    "LF2": {"levels": 6, "boxes": [[-0.5, 2.0, -0.75, 0.75, 4], [1.0, 4.0, -0.35, 0.65, 3], [-2.0, 6.0, -3.0, 3.0, 1],
                                   [2.0, 3.25, -0.35, 0.65, 4]]},
    "L7": {"levels": 7, "boxes": [[-0.5, 2.0, -0.75, 0.75, 5], [1.0, 4.0, -0.35, 0.65, 4], [-2.0, 6.0, -3.0, 3.0, 2],
                                  [2.0, 3.25, -0.35, 0.65, 5]]},
    "L8": {"levels": 8, "boxes": [[-0.5, 2.0, -0.75, 0.75, 6], [1.0, 4.0, -0.35, 0.65, 5], [-2.0, 6.0, -3.0, 3.0, 3],
                                  [2.0, 3.25, -0.35, 0.65, 6]]},
}

# Viscous time-step limit (reference AmrViscousStabilization, src/static_amr2d_viscous.h:14, and
# the static-AMR command line's --viscous-stabilization, src/static_amr2d_cli.cpp:113): "srd" maps
# the diffusion rate through the SRD neighbourhoods of a non-agglomerated mesh.
STABILIZATION = {"volume": 0, "srd": 1}

BC_KINDS = {"transmissive": 1, "reflect": 2, "farfield": 3, "total-inlet": 4, "pressure-outlet": 5, "no-slip": 6}


def merged(overrides: dict | None = None, base: dict | None = None) -> dict:
    """Deep-merge `overrides` into a copy of the defaults; unknown keys are an error
    (pipeline/configuration.py:13-23)."""
    out = copy.deepcopy(base if base is not None else DEFAULTS)

    def walk(dst, src, path):
        for key, value in src.items():
            if key not in dst:
                raise KeyError(f"unknown configuration key {path + key}")
            if isinstance(value, dict) and isinstance(dst[key], dict):
                walk(dst[key], value, path + key + ".")
            else:
                dst[key] = copy.deepcopy(value)

    walk(out, overrides or {}, "")
    return out


def scales(c: dict) -> dict:
    """Reference scales (pipeline/configuration.py:129-140)."""
    f, p = c["freestream"], c["physics"]
    rho = f["p_Pa"] / (f["R_J_kgK"] * f["T_K"])
    a = math.sqrt(f["gamma"] * f["R_J_kgK"] * f["T_K"])
    law = p["sutherland"]
    mu = law["mu0_Pa_s"] * (f["T_K"] / law["T0_K"]) ** 1.5 * (law["T0_K"] + law["S_K"]) / (f["T_K"] + law["S_K"])
    if p["viscosity"] == "constant" and p["constant_mu_Pa_s"] is not None:
        mu = p["constant_mu_Pa_s"]
    return dict(rho=rho, velocity=a, length=f["chord_m"], pressure=rho * a * a, temperature=f["gamma"] * f["T_K"],
                time=f["chord_m"] / a, viscosity=rho * a * f["chord_m"], mu_inf=mu,
                q_inf=.5 * rho * (f["Mach"] * a) ** 2, Re_chord=rho * f["Mach"] * a * f["chord_m"] / mu)


@dataclass
class Case:
    """Nondimensional inputs of one solve (what the reference's settings.txt carries)."""
    gamma: float
    mach: float
    aoa: float
    cfl: float
    dt_fixed: float
    flux: int
    model: int
    sa_model: int
    law: int
    mu_star: float
    sutherland_star: float
    sa_nu_inf: float
    viscous_cfl: float
    sa_cfl: float
    wall_probe_min_h: float
    outlet_p_star: float
    bc_kinds: list
    mesh_options: dict
    convergence: dict = field(default_factory=dict)
    moment_reference: tuple = (0.25, 0.0)
    stop_time: float = 100.0
    stabilization: int = 0  # viscous dt limit: 0 true volume, 1 SRD-mapped diffusion rate

    @property
    def freestream(self) -> tuple:
        a = math.radians(self.aoa)
        return (1.0, self.mach * math.cos(a), self.mach * math.sin(a), 1.0 / self.gamma)

    @property
    def q_inf(self) -> float:
        return .5 * self.mach * self.mach


def make_case(preset: str = "LF2", overrides: dict | None = None) -> Case:
    """Nondimensionalise a configuration (pipeline/configuration.py:143-167 and the setup in
    src/quick_airfoil.cpp:35-48)."""
    c = merged(overrides)
    f, p, s, m = c["freestream"], c["physics"], c["solver"], c["mesh"]
    sc = scales(c)
    mu = sc["mu_inf"] / sc["viscosity"]
    profile = PRESETS[preset]
    b = m["bounds_xy_c"]
    fixed = s["time_step"]["dt_seconds"] / sc["time"] if s["time_step"]["mode"] == "fixed" else 0.0
    model = 0 if p["model"] == "euler" else 1
    return Case(
        gamma=f["gamma"], mach=f["Mach"], aoa=f["AOA_degrees"], cfl=s["CFL"], dt_fixed=fixed,
        flux=0 if s["flux"] == "hllc" else 1, model=model,
        sa_model=1 if (model == 1 and p["turbulence"] == "sa-neg") else 0,
        law=1 if p["viscosity"] == "sutherland" else 0, mu_star=mu,
        sutherland_star=p["sutherland"]["S_K"] / sc["temperature"], sa_nu_inf=p["sa_nu_inf_over_nu"] * mu,
        viscous_cfl=s["viscous_CFL"], sa_cfl=s["SA_CFL"], wall_probe_min_h=p["wall_probe_min_h"],
        outlet_p_star=c["boundaries"]["outlet_pressure_Pa"] / sc["pressure"],
        bc_kinds=[BC_KINDS[c["boundaries"][side]] for side in ("xmin", "xmax", "ymin", "ymax")],
        mesh_options=dict(nx=m["base_shape"][0], ny=m["base_shape"][1], levels=profile["levels"],
                          buffer=m["wall_buffer"], xmin=b[0], xmax=b[1], ymin=b[2], ymax=b[3],
                          max_cells=m["max_cells"], agglomerate_small=bool(m["agglomerate_small"]),
                          boxes=[list(x) for x in profile["boxes"]]),
        convergence=dict(s["convergence"]), moment_reference=tuple(s["moment_reference_xy_c"]),
        stop_time=s["stop_time_star"], stabilization=STABILIZATION[p["viscous_stabilization"]])
