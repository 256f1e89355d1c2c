"""amr_hip_run_noalloc (the solve that allocates nothing) against the session path it restates:
amr_hip_create + amr_hip_advance + amr_hip_download, bit for bit (same kernels, same launch
tile; graph replay and direct launch give the same bits).  Two calls in a row, the second on a
smaller mesh, so the second runs in the same static device array the first left dirty; a mesh
beyond the array's 64 MiB is rejected with a message; a fatal run writes back the last accepted
state as the session does."""
import ctypes as C

import numpy as np
import pytest

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import _native as N
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import BackendError, Session

from .common import LAMINAR, case_mesh

pytestmark = pytest.mark.gpu


def run_noalloc(mesh, params, u0, tend, precision, maxsteps=50_000_000, budget=0.0):
    lib = N.load_library()
    u = np.array(u0, dtype=np.float64, copy=True)
    meta = N.amr_state_meta(0.0, 0, 0)
    err = N.err_buffer()
    view = mesh.view
    status = lib.amr_hip_run_noalloc(C.byref(view), C.byref(params), N.dptr(u), u.size, C.byref(meta), tend,
                                     maxsteps, budget, precision, err, N.ERR_LEN)
    return status, u, dict(time=meta.time, steps=meta.steps, rejected=meta.rejected), err.value.decode()


def session_run(mesh, params, u0, tend, precision, maxsteps=50_000_000):
    with Session(mesh, params, u0, precision=precision) as s:
        try:
            status = s.advance(tend, maxsteps=maxsteps)
            message = ""
        except BackendError as e:
            status, message = 1, str(e)
        return status, s.download(), dict(s.meta), message


@pytest.mark.parametrize("precision", [64, 32])
def test_noalloc_matches_session_on_two_calls(precision):
    for preset, tend in (("LF2", 0.02), ("LF1", 0.01)):  # the second call is smaller
        case, mesh, params = case_mesh("G0001", preset)
        u0 = mesh.initialize(params, case.freestream)
        print("noalloc", preset, mesh.n_cells, "cells, precision", precision, flush=True)
        status, u, meta, message = run_noalloc(mesh, params, u0, tend, precision)
        ref_status, ref_u, ref_meta, _ = session_run(mesh, params, u0, tend, precision)
        print(preset, status, meta, ref_meta, message, flush=True)
        assert status == ref_status == 0 and meta == ref_meta
        assert np.array_equal(u, ref_u)


def test_noalloc_fatal_and_limits():
    # A step cap (fatal 3) writes back the last accepted state, as the session path does.
    case, mesh, params = case_mesh("G0002", "LF1", LAMINAR)
    u0 = mesh.initialize(params, case.freestream)
    status, u, meta, message = run_noalloc(mesh, params, u0, 0.01, 64, maxsteps=7)
    ref_status, ref_u, ref_meta, ref_message = session_run(mesh, params, u0, 0.01, 64, maxsteps=7)
    print("fatal", status, meta, message, ref_meta, flush=True)
    assert status == ref_status == 1 and meta == ref_meta and meta["steps"] == 7
    assert message == ref_message and "status 3" in message
    assert np.array_equal(u, ref_u)
    # Invalid arguments are rejected before any device work.
    status, _, _, message = run_noalloc(mesh, params, u0, -1.0, 64)
    assert status == 1 and message == "invalid AMR GPU advance"
    # A level-8 mesh (about 100k cells) needs more than the 64 MiB array in FP64.
    case, big, bparams = case_mesh("G0001", "L8")
    print("large mesh", big.n_cells, "cells", flush=True)
    status, _, _, message = run_noalloc(big, bparams, big.initialize(bparams, case.freestream), 0.01, 64)
    print(message, flush=True)
    assert status == 1 and "the device scratch holds 67108864" in message
