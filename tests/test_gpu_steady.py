"""Full steady solves on the GPU against the reference's published results.

The reference publishes, for the G0001 example airfoil with the default configuration (SA-neg,
Sutherland, HLLC, CFL 0.15, viscous/SA CFL 0.1, checks every t* = 0.5, hold 5 after t* = 3),
the steps and time to steady and CL / CD / CM_z on LF1 and LF2 (validation/summary.json,
full_steady).  The FP64 device path (session + on-device monitor + check loop) must reproduce
the step count and the stop time exactly and the coefficients within 1e-8; the FP32 mode must
stay within 1e-4 of them.
"""
import pytest

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import Session

from .common import case_mesh, pins

pytestmark = pytest.mark.gpu


def published(profile):
    return {r["profile"]: r for r in pins()["full_steady_G0001"]}[profile]


@pytest.mark.parametrize("profile", ["LF1", "LF2"])
def test_fp64_steady_matches_published(profile):
    pub = published(profile)
    case, mesh, params = case_mesh("G0001", profile)
    assert mesh.n_cells == pub["cells"]
    with Session(mesh, params, mesh.initialize(params, case.freestream)) as s:
        r = s.solve(case, mode="steady", budget=100, history=0)
    print(profile, r["steps"], r["time"], r["cl"], r["cd"], r["cm"], r["stop_reason"], flush=True)
    assert r["converged"] and not r["failed"]
    assert r["steps"] == pub["steps"] and r["time"] == pub["time_star"]
    for q in ("cl", "cd", "cm"):
        assert abs(r[q] - pub[q]) <= 1e-8, q


def test_fp32_steady_within_1e4_of_published():
    pub = published("LF2")
    case, mesh, params = case_mesh("G0001", "LF2")
    with Session(mesh, params, mesh.initialize(params, case.freestream), precision=32) as s:
        r = s.solve(case, mode="steady", budget=100, history=0)
    print("fp32", r["steps"], r["time"], r["cl"], r["cd"], r["cm"], r["stop_reason"], flush=True)
    assert r["converged"]
    for q in ("cl", "cd"):
        assert abs(r[q] - pub[q]) <= 1e-4, q
    assert abs(r["cm"] - pub["cm"]) <= 1e-4
