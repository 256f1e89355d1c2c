"""The legacy-grid GPU backend API (include/euler2d_gpu.h: cfd::solver2d_run_cuda and
cfd::solver2d_cuda_available) as a caller compiled against it sees it (tests/capi/ugrid_caller).

CPU: the Solver2D layout of the public header, of the library's private view and of the
reference's own type (oracle/_ref) agree field for field; without a device the API reports
unavailable and returns 3.  GPU: a run through the C++ API is bitwise the C ABI session's on the
same initial state (regular grids, and a Cut2D filled by the caller for SSP-RK3 and backward
Euler), and refusals return 3.
"""
import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import pygrid as G
from oracle import pyoracle as O
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import _native as N
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import euler2d as E

from .grid_cases import params

CALLER = Path(__file__).resolve().parent / "capi" / "ugrid_caller"
needs_caller = pytest.mark.skipif(not CALLER.exists(), reason="tests/capi/ugrid_caller not built")


def library_layout():
    out = (C.c_int64 * 64)()
    n = N.load_library().ugrid_solver2d_layout(out, 64)
    return list(out[:n])


@needs_caller
def test_layouts_agree():
    header = [int(x) for x in subprocess.run([str(CALLER), "layout"], capture_output=True, text=True,
                                             check=True).stdout.split()]
    assert len(header) == 50
    assert header == library_layout()
    lib = O.ref_lib()
    if lib is None:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    lib.ref_solver2d_layout.restype = C.c_int
    lib.ref_solver2d_layout.argtypes = [C.POINTER(C.c_int64), C.c_int32]
    out = (C.c_int64 * 64)()
    n = lib.ref_solver2d_layout(out, 64)
    assert list(out[:n]) == header


@needs_caller
def test_unavailable_without_device():
    ok, _ = E.solver2d_cuda_available()
    if ok:
        pytest.skip("a device is present")
    probe = subprocess.run([str(CALLER), "probe"], capture_output=True, text=True, check=True).stdout
    assert probe.startswith("0 ")
    run = subprocess.run([str(CALLER), "run", "/dev/null", "8", "2", "1", "0", "1", "1", "1", "1", "0", "0", "0.1",
                          "10"], capture_output=True, text=True, check=True)
    assert "status 3" in run.stdout and "unavailable" in run.stderr


def caller_run(tmp_path, nx, ny, ncase, flux, bc, stretch, tend, maxsteps, limiter=3, radius=0.0, be=0):
    out = tmp_path / "run.bin"
    args = [str(CALLER), "run", str(out), str(nx), str(ny), str(ncase), str(flux), *map(str, bc), str(stretch[0]),
            str(stretch[1]), repr(tend), str(maxsteps), str(limiter), repr(radius), str(be)]
    print(" ".join(args[2:]), flush=True)
    done = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert done.returncode == 0, done.stderr
    raw = out.read_bytes()
    head = np.frombuffer(raw[:24], dtype=np.int64)
    body = np.frombuffer(raw[24:], dtype=np.float64)
    states = ncase * 4 * (nx + 4) * (ny + 4)
    return dict(status=int(head[0]), nsteps=int(head[1]), nreject=int(head[2]), initial=body[:states].copy(),
                u=body[states:2 * states].copy(), t=body[2 * states:2 * states + ncase].copy(),
                dt=body[2 * states + ncase:].copy(), stderr=done.stderr)


@pytest.mark.gpu
@needs_caller
@pytest.mark.parametrize("flux, bc, stretch", [
    (E.FLUX_HLLC, (1, 1, 1, 1), (0, 0.0)),
    (E.FLUX_AUSMPUP, (2, 1, 0, 0), (1, 1.01)),
    (E.FLUX_HLLC, (2, 2, 2, 2), (3, 1.5)),
])
def test_cpp_api_run_bitwise(tmp_path, flux, bc, stretch):
    nx, ny, ncase, tend = 150, 20, 4, 0.04
    res = caller_run(tmp_path, nx, ny, ncase, flux, bc, stretch, tend, 100000)
    assert res["status"] == 0 and res["nsteps"] > 10, res["stderr"]
    s = E.Solver2D(nx, ny, 0.0, 1.0, 0.0, 1.0, ncase, params(flux=flux, bc=bc, sx=stretch))
    s.U[:] = res["initial"]
    s.mco[:] = [0.2 + 0.2 * c for c in range(ncase)]
    ora = G.run(s, tend, 100000)
    assert E.solver2d_run_cuda(s, tend, 100000) == 0
    np.testing.assert_array_equal(res["u"], s.U)
    np.testing.assert_array_equal(res["t"], s.t)
    np.testing.assert_array_equal(res["dt"], s.dt)
    assert (res["nsteps"], res["nreject"]) == (s.nsteps, s.nreject) == (ora["nsteps"], ora["nreject"])
    np.testing.assert_array_equal(res["u"], ora["u"])
    assert "hip: device=" in res["stderr"]


@pytest.mark.gpu
@needs_caller
def test_cpp_api_refusals(tmp_path):
    res = caller_run(tmp_path, 16, 4, 1, 0, (1, 1, 1, 1), (0, 0.0), 0.1, 10, limiter=2)
    assert res["status"] == 3 and "requires --limiter first" in res["stderr"]
    np.testing.assert_array_equal(res["u"], res["initial"])
    res = caller_run(tmp_path, 16, 4, 1, 0, (1, 1, 1, 1), (0, 0.0), 0.1, 0)
    assert res["status"] == 1 and res["nsteps"] == 0


@pytest.mark.gpu
@needs_caller
@pytest.mark.parametrize("be, nx, ny, tend", [(0, 120, 90, 0.05), (1, 64, 48, 0.02)])
def test_cpp_api_cut_run_bitwise(tmp_path, be, nx, ny, tend):
    ncase, radius = 2, 0.21
    res = caller_run(tmp_path, nx, ny, ncase, E.FLUX_HLLC, (1, 1, 2, 2), (0, 0.0), tend, 100000, radius=radius, be=be)
    assert res["status"] == 0 and res["nsteps"] > 1, res["stderr"]
    par = params(bc=(1, 1, 2, 2))
    if be:
        par.time_integrator = E.TIME_BACKWARD_EULER
        par.cfl = 2.0
    s = E.Solver2D(nx, ny, 0.0, 1.0, 0.0, 1.0, ncase, par)
    s.U[:] = res["initial"]
    s.mco[:] = [0.2 + 0.2 * c for c in range(ncase)]
    # the caller's disc, vertex for vertex (libm's cos / sin, as the caller's)
    a = [0.0131 + 2.0 * 3.141592653589793 * k / 48 for k in range(48)]
    s.set_cut(E.cut2d_build_polygon(s, [0.47 + radius * math.cos(x) for x in a],
                                    [0.53 + radius * math.sin(x) for x in a]))
    assert s.cut.nsmall > 0
    ora = (G.run_implicit if be else G.run)(s, tend, 100000)
    np.testing.assert_array_equal(res["u"], ora["u"])
    np.testing.assert_array_equal(res["t"], ora["t"])
    np.testing.assert_array_equal(res["dt"], ora["dt"])
    assert res["nsteps"] == ora["nsteps"]


def test_cut_setup_refusals_before_any_device_work():
    # ugrid_create validates the cut geometry before it touches a device
    from .grid_cases import cut_body
    s, px, py = cut_body("circle")
    g = s.grid
    k = (g.my - 3) * g.mx + g.mx // 2  # an interior cell two rows below the top side
    s.cut.flag[k] = 1.0
    with pytest.raises(RuntimeError, match="within three cells of the domain boundary"):
        E.DeviceSession(s)
    s.cut.flag[k] = 0.0
    s.cut.vol[2] = 0.0  # a ghost cell
    with pytest.raises(RuntimeError, match="within three cells"):
        E.DeviceSession(s)
    s.cut.vol[2] = s.grid.dx[2] * s.grid.dy[0]
    s.cut.smem[0] = float(g.mx * g.my)
    with pytest.raises(RuntimeError, match="invalid SRD member"):
        E.DeviceSession(s)
    s.cut.nbad_nbhd = 1
    with pytest.raises(RuntimeError, match="refuses 1 invalid SRD neighbourhoods"):
        E.DeviceSession(s)


def test_cut_builder_memory_safety_under_sanitizers(tmp_path):
    """The native cut-cell builder (host code) compiled with AddressSanitizer and
    UndefinedBehaviorSanitizer: every view array read end to end on three grids and three bodies,
    and the refusals (too few vertices, a body at the sides, an edge on a grid line) clean."""
    import shutil
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    root = Path(__file__).resolve().parents[1]
    host = root / "quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd" / "csrc" / "host"
    exe = tmp_path / "cut_builder_asan"
    build = subprocess.run([gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                            "-I", str(root / "include"), "-I", str(host), str(CALLER.parent / "cut_builder_asan.cpp"),
                            str(host / "ugrid_cut.cpp"), str(host / "cutcell.cpp"), str(host / "ugrid_axis.cpp"),
                            "-o", str(exe)], capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300,
                         env={"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "halt_on_error=1"})
    assert run.returncode == 0, run.stdout + run.stderr
    assert "all ok" in run.stdout and run.stdout.count("ok nx=") == 9
