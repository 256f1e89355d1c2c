"""The drop-in C++ API (include/static_amr2d_gpu.h) driven by a compiled caller (tests/capi/api_caller.cpp)
that holds a cfd::StaticAmrMesh2D and a cfd::StaticAmrState2D the way the reference's solver does
(src/quick_airfoil.cpp:61-97).

GPU: create / advance / download / stats / destroy give exactly the bits and counters of the Python
session over the C ABI (which is pinned to the oracle); a wall budget returns 2 with a resumable
state; invalid configurations return 1 with the reference's messages.  CPU: without a device,
create returns 1 with a message and nothing falls back to the host (the reference stub's
behaviour, src/static_amr2d_gpu_stub.cpp:6-12).
"""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import session as S
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import Session

from .common import case_mesh

CALLER = Path(__file__).resolve().parent / "capi" / "api_caller"


def write_inputs(tmp: Path, mesh, params, u0, t1, t2):
    arrays = mesh.arrays()
    names = ["cells", "faces", "walls", "face_offsets", "cell_faces", "group_offsets", "group_members",
             "member_offsets", "member_groups", "group_volume", "visc_offsets", "visc_terms", "wall_samples",
             "face_samples", "face_geometry", "wall_distance"]
    for name in names:
        if name in arrays:
            (tmp / f"{name}.bin").write_bytes(np.ascontiguousarray(arrays[name]).tobytes())
    view = mesh.view
    (tmp / "boundaries.bin").write_bytes(bytes(view.boundaries))
    (tmp / "transport.bin").write_bytes(bytes(view.transport))
    (tmp / "params.bin").write_bytes(bytes(params))
    (tmp / "u0.bin").write_bytes(np.ascontiguousarray(u0, dtype=np.float64).tobytes())
    (tmp / "scalars.txt").write_text(f"{int(view.small_cells)} {int(view.agglomerate_small)} {t1!r} {t2!r}\n")


def params_as_reference(p):
    """The bytes of a cfd::Params2D (src/euler2d.h:38-52, 136 bytes) holding the path's fields of
    `p`, the rest at the reference's defaults (params2d_default, src/euler2d.cpp:13-37)."""
    raw = struct.pack("<2d2i2d4idi4xdi4x2di4xdi4xd", p.gamma, p.cfl, p.flux, p.limiter, 5.0, p.ausm_mco, *p.bc,
                      p.dt_fixed, p.time_integrator, 0.4, 2000, 1e-8, 1e-11, 0, 0.0, 0, 0.0)
    assert len(raw) == 136
    return raw


class Raw:
    def __init__(self, data):
        self.data = data

    def __bytes__(self):
        return self.data


def run_caller(tmp):
    done = subprocess.run([str(CALLER), str(tmp)], capture_output=True, text=True, timeout=300)
    lines = {}
    if (tmp / "result.txt").exists():
        for line in (tmp / "result.txt").read_text().splitlines():
            key, _, rest = line.partition(" ")
            lines[key] = rest
    return done, lines


@pytest.mark.skipif(not CALLER.exists(), reason="api_caller not built")
def test_cpp_api_without_device_fails_loudly(tmp_path):
    if S.device_count() > 0:
        pytest.skip("a HIP device is visible")
    case, mesh, params = case_mesh("G0001", "LF1")
    write_inputs(tmp_path, mesh, Raw(params_as_reference(params)), mesh.initialize(params, case.freestream), 0.002,
                 0.004)
    done, lines = run_caller(tmp_path)
    assert done.returncode == 1 and lines["create"] == "1"
    assert "no HIP device" in done.stderr


@pytest.mark.gpu
def test_cpp_api_matches_session_bitwise(tmp_path):
    case, mesh, params = case_mesh("G0001", "LF1")
    u0 = mesh.initialize(params, case.freestream)
    t1, t2 = 0.01, 0.02
    write_inputs(tmp_path, mesh, Raw(params_as_reference(params)), u0, t1, t2)
    print("api_caller on G0001 LF1 SA-neg", mesh.n_cells, "cells", flush=True)
    done, lines = run_caller(tmp_path)
    print(done.stdout, done.stderr, lines, flush=True)
    assert done.returncode == 0
    with Session(mesh, params, u0) as s:
        assert s.advance(t1) == 0
        m1 = dict(s.meta)
        u1 = s.download()
        assert s.advance(t2) == 0
        m2 = dict(s.meta)
        u2 = s.download()
    got = np.frombuffer((tmp_path / "u.bin").read_bytes(), dtype=np.float64)
    assert np.array_equal(got, u2)
    adv = [row for row in (tmp_path / "result.txt").read_text().splitlines() if row.startswith("advance ")]
    for row, m in zip(adv, (m1, m2)):
        status, down, time, steps, rejected = row.split()[1:]
        assert (int(status), int(down)) == (0, 0)
        assert (float(time), int(steps), int(rejected)) == (m["time"], m["steps"], m["rejected"])
    polls, downloads, d2h, allocated, nodes = map(int, lines["stats"].split())
    assert polls >= 2 and downloads == 2 and d2h == 2 * u0.size * 8 and allocated > 0 and nodes >= 7
    # Wall budget: 2 after the first poll (32 of the 41 steps), then resumed to the same bits.
    b = lines["budget"].split()
    assert (int(b[0]), int(b[1])) == (0, 2) and float(b[2]) < t1 and int(b[3]) == 32
    assert (int(b[4]), int(b[5]), float(b[6]), int(b[7]), int(b[8])) == (0, 0, t1, m1["steps"], 1)
    assert np.array_equal(np.frombuffer((tmp_path / "u_resumed.bin").read_bytes(), dtype=np.float64), u1)
    assert lines["behind"] == "1 invalid AMR GPU advance"
    assert lines["cfl"] == "1 1 invalid AMR GPU configuration"
    assert lines["trials"].startswith("1 1 ") and len(lines["trials"]) > 4
    assert lines["size"] == "1 1 invalid AMR GPU configuration"
    assert lines["null"] == "0 1 1 invalid AMR GPU advance|invalid AMR GPU download"
