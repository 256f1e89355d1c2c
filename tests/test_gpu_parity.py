"""HIP path vs the FP64 CPU oracle (parity proper; needs the GPU).

Tolerances (FP64 path): one residual evaluation within 1e-12 relative L2 per variable (the two
differ only in operation order and FMA contraction); states after a t* = 0.01 window within
1e-10 relative L2 per variable with identical step and rejection counts.  FP32 path: states
within 2e-4 relative L2 over the same window (state and fluxes in FP32, control in FP64).
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import BackendError, Session

from .common import EULER, LAMINAR, case_mesh, merge, naca_polygon, per_variable_l2, pins

pytestmark = pytest.mark.gpu

SRD = {"mesh": {"agglomerate_small": False}}
AUSM = {"solver": {"flux": "ausm"}}

CASES = {
    "sa_hllc": (None,),
    "laminar_hllc": (LAMINAR,),
    "laminar_ausm": (merge(LAMINAR, AUSM),),
    "euler_hllc_agglomerated": (EULER,),
    "euler_hllc_srd": (merge(EULER, SRD),),
    "laminar_srd": (merge(LAMINAR, SRD),),
}


def perturbed(u, n, seed=3):
    rng = np.random.default_rng(seed)
    v = u.copy()
    v[:n] *= 1 + 0.02 * rng.standard_normal(n)
    v[n:2 * n] *= 1 + 0.05 * rng.standard_normal(n)
    v[2 * n:3 * n] += 0.05 * rng.standard_normal(n)
    return v


@pytest.mark.parametrize("name", list(CASES))
def test_residual_parity_fp64(name):
    overrides, = CASES[name]
    case, mesh, params = case_mesh("G0001", "LF1", overrides)
    u0 = mesh.initialize(params, case.freestream)
    u1 = perturbed(u0, mesh.n_cells)
    with Session(mesh, params, u0) as s:
        for u in (u0, u1):
            ref = O.rhs(mesh, params, u)
            got = s.rhs(u)
            errs = per_variable_l2(got, ref, s.nv)
            print(name, "residual relative L2 per variable", errs, flush=True)
            assert max(errs) <= 1e-12


@pytest.mark.parametrize("name", list(CASES))
def test_advance_parity_fp64(name):
    overrides, = CASES[name]
    case, mesh, params = case_mesh("G0001", "LF1", overrides)
    u0 = mesh.initialize(params, case.freestream)
    ref, meta = O.run(mesh, params, u0, 0.01)
    assert meta["status"] == 0, meta["error"]
    with Session(mesh, params, u0) as s:
        assert s.advance(0.01) == 0
        got = s.download()
        print(name, "steps", s.meta, meta, flush=True)
        assert s.meta["steps"] == meta["steps"] and s.meta["rejected"] == meta["rejected"]
        assert s.meta["time"] == meta["time"] == 0.01
    errs = per_variable_l2(got, ref, 5 if case.sa_model else 4)
    print(name, "state relative L2 per variable", errs, flush=True)
    assert max(errs) <= 1e-10


@pytest.mark.parametrize("preset", ["LF1", "LF2"])
def test_published_step_counts(preset):
    """G0001 default SA-neg case to t* = 0.01: 41 / 82 steps (validation/ten_short_checks.json)."""
    expected = {(r["case"], r["profile"]): r["steps"] for r in pins()["short_runs_t0.01"]}[("G0001", preset)]
    case, mesh, params = case_mesh("G0001", preset)
    u0 = mesh.initialize(params, case.freestream)
    with Session(mesh, params, u0) as s:
        s.advance(0.01)
        print(preset, s.meta, flush=True)
        assert s.meta["steps"] == expected and s.meta["rejected"] == 0


def test_euler_lf1_naca0012_srd():
    """The Euler LF1 case: symmetric 12% NACA-style section, Euler, SRD small cells, t* = 0.01."""
    case, mesh, params = case_mesh(naca_polygon(0.12), "LF1", merge(EULER, SRD))
    assert mesh.info.small_cells > 0
    u0 = mesh.initialize(params, case.freestream)
    ref, meta = O.run(mesh, params, u0, 0.01)
    with Session(mesh, params, u0) as s:
        s.advance(0.01)
        got = s.download()
        assert s.meta["steps"] == meta["steps"]
    assert max(per_variable_l2(got, ref, 4)) <= 1e-10


def test_fp32_window():
    case, mesh, params = case_mesh("G0001", "LF1")
    u0 = mesh.initialize(params, case.freestream)
    ref, meta = O.run(mesh, params, u0, 0.01)
    with Session(mesh, params, u0, precision=32) as s:
        s.advance(0.01)
        got = s.download()
        print("fp32 steps", s.meta, meta, flush=True)
        assert abs(s.meta["steps"] - meta["steps"]) <= 1
    errs = per_variable_l2(got, ref, 5)
    print("fp32 state relative L2", errs, flush=True)
    assert max(errs) <= 2e-4


def test_graph_and_direct_launch_identical():
    case, mesh, params = case_mesh("G0001", "LF1")
    u0 = mesh.initialize(params, case.freestream)
    out = []
    for graph in (True, False):
        with Session(mesh, params, u0, graph=graph, batch_trials=7) as s:
            s.advance(0.004)
            out.append((s.download(), dict(s.meta)))
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


def test_rest_state_stays_at_rest():
    """Zero-velocity freestream stays at rest (src/quick_airfoil.cpp:98-101, L_inf <= 2e-11)."""
    case, mesh, params = case_mesh("G0001", "LF1", LAMINAR)
    case.mach = 0.0  # reference "rest" mode: zero velocity inside and in the farfield state
    mesh.set_physics(case)
    mesh.prepare_viscous(params)
    u0 = mesh.initialize(params, case.freestream)
    with Session(mesh, params, u0) as s:
        s.advance(0.01)
        got = s.download()
    assert np.max(np.abs(got - u0)) <= 2e-11


def test_maxsteps_is_fatal_3():
    case, mesh, params = case_mesh("G0001", "LF1")
    u0 = mesh.initialize(params, case.freestream)
    with Session(mesh, params, u0) as s:
        with pytest.raises(BackendError, match="status 3"):
            s.advance(0.01, maxsteps=3)
        assert s.meta["steps"] == 3
        again = s.download()
    ref, meta = O.run(mesh, params, u0, 0.01, maxsteps=3)
    assert meta["steps"] == 3 and meta["status"] == 1
    assert max(per_variable_l2(again, ref, 5)) <= 1e-12


def test_budget_returns_2_and_resumes():
    case, mesh, params = case_mesh("G0001", "LF1")
    u0 = mesh.initialize(params, case.freestream)
    with Session(mesh, params, u0, batch_trials=1) as s:
        status = s.advance(0.5, wall_budget=1e-6)
        assert status == 2 and 0 < s.meta["steps"]
        assert s.advance(0.01 if s.meta["time"] < 0.01 else s.meta["time"]) == 0


def test_monitor_matches_oracle():
    """Device FP64 monitor vs the oracle's steady monitor (src/quick_airfoil.cpp:75-85) at the
    same accepted state: residual norms within 1e-10 relative, CL/CD/CM_z within 1e-10."""
    case, mesh, params = case_mesh("G0001", "LF1")
    u0 = mesh.initialize(params, case.freestream)
    region, moment = case.convergence["residual_region_xy_c"], case.moment_reference
    with Session(mesh, params, u0) as s:
        s.advance(0.01)
        u = s.download()
        got = s.monitor(region, moment, case.mach, case.aoa)
    ref = O.monitor(mesh, params, u, region, moment, case.mach, case.aoa)
    print("monitor", got, ref, flush=True)
    assert np.allclose(got["residual"], ref["residual"], rtol=1e-10, atol=0)
    for key in ("cl", "cd", "cm"):
        assert abs(got[key] - ref[key]) <= 1e-10


def test_time_driver_to_half_matches_oracle():
    """Device driver in time mode to t* = 0.5 (checks at 0.01 and 0.5), against the oracle run
    with the same two targets: equal steps, state within 1e-10 relative L2 per variable (measured
    about 1.5e-14 after 2083 steps), CL/CD/CM_z within 1e-10."""
    case, mesh, params = case_mesh("G0001", "LF1")
    u0 = mesh.initialize(params, case.freestream)
    with Session(mesh, params, u0) as s:
        res = s.solve(case, mode="time", stop_time=0.5)
        got = s.download()
    assert not res["failed"] and res["checks"] == 2 and res["time"] == 0.5
    ref, meta = O.run(mesh, params, u0, 0.01)
    ref, meta2 = O.run(mesh, params, ref, 0.5, time=meta["time"], steps=meta["steps"])
    assert res["steps"] == meta2["steps"]
    errs = per_variable_l2(got, ref, 5)
    mon = O.monitor(mesh, params, ref, case.convergence["residual_region_xy_c"], case.moment_reference, case.mach,
                    case.aoa)
    print("time driver", res["steps"], errs, res["cl"], mon["cl"], flush=True)
    assert max(errs) <= 1e-10
    for key in ("cl", "cd", "cm"):
        assert abs(res[key] - mon[key]) <= 1e-10
