"""The C ABI library loads without a GPU and exports every entry point include/amr_hip.h and
include/static_amr2d_gpu.h declare; without a device the backend fails loudly."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import _native as N
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import BackendError, Session

from .common import case_mesh

ROOT = Path(__file__).resolve().parents[1]


def declared_c_symbols():
    text = (ROOT / "include" / "amr_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|int32_t|uint64_t)\s+(amr_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    lib = N.load_library()
    names = declared_c_symbols()
    assert len(names) >= 45
    for name in names:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(N.LIBRARY_PATH)], capture_output=True, text=True).stdout
    for fn in ("static_amr_gpu_create", "static_amr_gpu_advance", "static_amr_gpu_download", "static_amr_gpu_stats",
               "static_amr_gpu_destroy"):
        assert re.search(r"_ZN3cfd\d+" + fn, out), fn


def test_record_layouts_match_reference_sizes():
    assert C.sizeof(N.amr_cell) == 88 and C.sizeof(N.amr_face) == 56 and C.sizeof(N.amr_wall) == 80
    assert C.sizeof(N.amr_grad_term) == 24 and C.sizeof(N.amr_wall_sample) == 16
    assert C.sizeof(N.amr_face_geometry) == 40 and C.sizeof(N.amr_boundary) == 104
    assert C.sizeof(N.amr_transport) == 88


def test_create_without_device_fails_loudly():
    if N.load_library().amr_hip_device_count() > 0:
        pytest.skip("a HIP device is visible")
    case, mesh, params = case_mesh("G0001", "LF1")
    u0 = mesh.initialize(params, case.freestream)
    with pytest.raises(BackendError, match="no HIP device"):
        Session(mesh, params, u0)


def test_create_rejects_invalid_configuration():
    case, mesh, params = case_mesh("G0001", "LF1")
    u0 = mesh.initialize(params, case.freestream)
    params.cfl = 0.7
    with pytest.raises(BackendError, match="invalid AMR GPU configuration"):
        Session(mesh, params, u0)
    params.cfl = 0.15
    with pytest.raises(BackendError, match="invalid AMR GPU configuration"):
        Session(mesh, params, u0[:-1])


def check_layout(items, tile):
    lib = N.load_library()
    views = (N.amr_mesh_view * len(items))(*[m.view for _, m, _ in items])
    params = (N.amr_params * len(items))(*[p for _, _, p in items])
    err = N.err_buffer()
    return lib.amr_hip_check_layout(views, params, len(items), tile, err, N.ERR_LEN), err.value.decode()


@pytest.mark.parametrize("tile", [64, 128, 256])
def test_device_layout_is_consistent(tile):
    """Host-side audit of the device layout (storage order, slot ranges, CSR entries, gradient
    terms, SRD tables, wall owners/donors) for a mixed set of meshes at every launch tile."""
    from .common import EULER, LAMINAR, merge, naca_polygon
    items = [case_mesh("G0001", "LF1"), case_mesh("G0002", "LF1", LAMINAR),
             case_mesh(naca_polygon(0.12), "LF1", merge(EULER, {"mesh": {"agglomerate_small": False}}))]
    status, message = check_layout(items[:1], tile)
    assert status == 0, message
    status, message = check_layout(items[1:], tile)
    assert status == 0, message
    # Four slots: three trial branches, the first with two slots (per-branch block tables, and
    # every slot's tile count in each, which the last stage pass's commit counts on).
    status, message = check_layout(items[1:] + [case_mesh("G0100", "LF1", LAMINAR), case_mesh("G0010", "LF1", LAMINAR)],
                                   tile)
    assert status == 0, message
