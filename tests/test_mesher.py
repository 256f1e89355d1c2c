"""Host mesher: bit-identical to the reference mesher; published cell counts; geometry audit.

Pins: the 20 (airfoil, level) fluid cell counts of validation/ten_short_checks.json; the
geometric audit of src/quick_airfoil.cpp:51-58 (area, wall length, closure); and, where the
reference CPU code is built (oracle/_ref, build container), every mesh and viscous-table array
compared field by field with the reference mesher's output, at LF1 / LF2 and at levels 7 and 8.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import geometry as G

from .common import AIRFOILS, EULER, case_mesh, merge, naca_polygon, pins

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference CPU build (oracle/_ref) not present")


def same_records(a, b) -> bool:
    if a.dtype.names:
        return all(np.array_equal(a[f], b[f]) and a[f].tobytes() == b[f].tobytes() for f in a.dtype.names)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("profile", ["LF1", "LF2"])
@pytest.mark.parametrize("name", AIRFOILS)
def test_published_cell_counts(name, profile):
    expected = {(r["case"], r["profile"]): r["cells"] for r in pins()["short_runs_t0.01"]}[(name, profile)]
    _, mesh, _ = case_mesh(name, profile)
    assert mesh.n_cells == expected


@needs_ref
@pytest.mark.parametrize("profile", ["LF1", "LF2"])
@pytest.mark.parametrize("name", ["G0001", "G0300", "G0480"])
def test_bit_identical_to_reference_sa(name, profile):
    _, mine, _ = case_mesh(name, profile)
    _, ref, _ = case_mesh(name, profile, lib=O.ref_lib(), prefix="ref_mesh")
    a, b = mine.arrays(), ref.arrays()
    assert set(a) == set(b)
    bad = [k for k in b if not same_records(a[k], b[k])]
    assert not bad, bad


@needs_ref
@pytest.mark.parametrize("thickness", [0.12, 0.06])
def test_bit_identical_to_reference_euler_srd(thickness):
    """Non-agglomerated (SRD) meshes of symmetric sections: exercises small-cell groups."""
    over = merge(EULER, {"mesh": {"agglomerate_small": False}})
    _, mine, _ = case_mesh(naca_polygon(thickness), "LF1", over)
    _, ref, _ = case_mesh(naca_polygon(thickness), "LF1", over, lib=O.ref_lib(), prefix="ref_mesh")
    a, b = mine.arrays(), ref.arrays()
    assert mine.info.small_cells == ref.info.small_cells > 0
    bad = [k for k in b if not same_records(a[k], b[k])]
    assert not bad, bad


@needs_ref
@pytest.mark.parametrize("preset, over", [
    ("L7", None),  # the serving loop's level-7 meshes (SA-neg)
    ("L8", {"freestream": {"Mach": 0.75, "AOA_degrees": 2.0},  # the level-8 laminar AUSM+-up case
            "physics": {"turbulence": "none", "viscosity": "constant"}, "solver": {"flux": "ausm"},
            "boundaries": {"xmax": "pressure-outlet"}}),
])
def test_bit_identical_to_reference_deep_levels(preset, over):
    """Level-7 and level-8 meshes (about 29k and 104k cells) of a cambered 14% section: every
    mesh and viscous-table array equal to the reference mesher's."""
    _, mine, _ = case_mesh(naca_polygon(0.14, 0.03, 0.4), preset, over)
    _, ref, _ = case_mesh(naca_polygon(0.14, 0.03, 0.4), preset, over, lib=O.ref_lib(), prefix="ref_mesh")
    assert mine.n_cells == ref.n_cells > (100_000 if preset == "L8" else 25_000)
    a, b = mine.arrays(), ref.arrays()
    assert set(a) == set(b)
    bad = [k for k in b if not same_records(a[k], b[k])]
    assert not bad, bad


@pytest.mark.parametrize("name", ["G0001", "G0450"])
def test_geometry_audit(name):
    """Fluid area = domain - body area (1e-10), wall length = perimeter (1e-11), closure (1e-9)."""
    case, mesh, _ = case_mesh(name, "LF1")
    arr = mesh.arrays()
    poly, _ = G.reorder_points(G.read_points(__import__("tests.common", fromlist=["GOLDEN"]).GOLDEN
                                             / "airfoils" / f"{name}_airfoil_points.csv"))
    o = case.mesh_options
    domain = (o["xmax"] - o["xmin"]) * (o["ymax"] - o["ymin"])
    assert abs(arr["cells"]["volume"].sum() - (domain - abs(G.polygon_area(poly)))) <= 1e-10
    assert abs(arr["walls"]["length"].sum() - G.polygon_perimeter(poly)) <= 1e-11
    n = mesh.n_cells
    closure = np.zeros((n, 2))
    surface = np.zeros(n)
    f = arr["faces"]
    for side, sign in (("left", 1.0), ("right", -1.0)):
        ok = f[side] >= 0
        np.add.at(closure, (f[side][ok], f["axis"][ok]), sign * f["length"][ok])
        np.add.at(surface, f[side][ok], f["length"][ok])
    w = arr["walls"]
    np.add.at(closure, (w["owner_cell"].astype(np.int64), 0), w["nx"] * w["length"])
    np.add.at(closure, (w["owner_cell"].astype(np.int64), 1), w["ny"] * w["length"])
    np.add.at(surface, w["owner_cell"].astype(np.int64), w["length"])
    assert np.max(np.hypot(closure[:, 0], closure[:, 1]) / surface) <= 1e-9
    assert (arr["cells"]["volume"] > 0).all()


def test_csr_is_consistent():
    _, mesh, _ = case_mesh("G0001", "LF1")
    a = mesh.arrays()
    off, cf, faces = a["face_offsets"].astype(np.int64), a["cell_faces"].astype(np.int64), a["faces"]
    for k in range(0, mesh.n_cells, 97):
        for q in cf[off[k]:off[k + 1]]:
            assert k in (faces["left"][q], faces["right"][q])
    assert off[-1] == (faces["left"] >= 0).sum() + (faces["right"] >= 0).sum()


def test_invalid_inputs_raise():
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.mesh import MeshError
    poly, patches = naca_polygon()
    with pytest.raises(MeshError, match="three finest cells"):
        case_mesh((poly * 30.0, patches), "LF1")
    with pytest.raises(MeshError, match="invalid mesh options"):
        case_mesh((poly, patches), "LF1", {"mesh": {"wall_buffer": 1}})
