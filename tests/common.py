"""Shared helpers of the test suite: fixture airfoils, cases and mesh construction."""
from __future__ import annotations

import json
from functools import lru_cache
from pathlib import Path

import numpy as np

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import config as CF
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import geometry as G
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import mesh as M

GOLDEN = Path(__file__).resolve().parent / "golden"
AIRFOILS = sorted(p.stem.split("_")[0] for p in (GOLDEN / "airfoils").glob("G*_airfoil_points.csv"))

EULER = {"physics": {"model": "euler", "turbulence": "none"}}
LAMINAR = {"physics": {"turbulence": "none"}}


def pins() -> dict:
    return json.loads((GOLDEN / "published_pins.json").read_text())


@lru_cache(maxsize=None)
def example_polygon(name: str):
    """GXXX example airfoil (fixture copy of the reference's examples/) as polygon + patches."""
    return G.reorder_points(G.read_points(GOLDEN / "airfoils" / f"{name}_airfoil_points.csv"))


def naca_polygon(thickness=0.12, camber=0.0, position=0.4, stations=100):
    return G.reorder_points(G.naca4_points(thickness, camber, position, stations))


def merge(*parts):
    out: dict = {}
    for p in parts:
        for k, v in (p or {}).items():
            if isinstance(v, dict):
                out.setdefault(k, {}).update(v)
            else:
                out[k] = v
    return out


def case_mesh(name_or_polygon, preset="LF1", overrides=None, lib=None, prefix="amr_mesh", sides=None):
    case = CF.make_case(preset, overrides)
    if isinstance(name_or_polygon, str):
        poly, patches = example_polygon(name_or_polygon)
    else:
        poly, patches = name_or_polygon
    mesh = M.build_case_mesh(case, poly, patches, lib=lib, prefix=prefix, sides=sides)
    return case, mesh, M.params_of(case)


def relative_l2(a, b) -> float:
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def per_variable_l2(a, b, nv) -> list:
    a, b = np.asarray(a).reshape(nv, -1), np.asarray(b).reshape(nv, -1)
    return [relative_l2(a[v], b[v]) for v in range(nv)]
