"""HIP path vs the FP64 oracle on every domain-side kind and every step-control branch (GPU).

The oracle's versions of these branches are pinned bit for bit to the reference's own CPU code
(test_oracle_controls.py); here the device must give the oracle's step and rejection counts,
fatal code and time exactly, and states within 1e-10 relative L2 per variable (FP64), residuals
within 1e-12.  Every failure is forced from the host (a huge fixed step, a negative pressure, a
time at which dt underflows, a step cap): each kernel still ends normally, the control record
carries the fatal code.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import _native as N
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import batch as B
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import BackendError, Session

from .common import LAMINAR, case_mesh, per_variable_l2, pins
from .variants import CONTROLS, FIXED_SCHEDULES, VARIANTS, VISCOUS_SRD

pytestmark = pytest.mark.gpu


def perturbed(u, n, seed=5):
    rng = np.random.default_rng(seed)
    v = u.copy()
    v[:n] *= 1 + 0.02 * rng.standard_normal(n)
    v[n:2 * n] += 0.05 * rng.standard_normal(n)
    v[2 * n:3 * n] += 0.05 * rng.standard_normal(n)
    return v


def variant(name):
    src, preset, over, sides = VARIANTS[name]
    case, mesh, params = case_mesh(src, preset, over, sides=sides)
    return case, mesh, params, mesh.initialize(params, case.freestream)


def nv_of(mesh):
    return 5 if mesh.view.transport.sa_model else 4


def check_run(name, got, meta, ref, ref_meta, nv, tol=1e-10, target=None):
    """Equal step and rejection counts; the time is equal bit for bit where a window ends at its
    target (the last step is clipped to it), otherwise it is a sum of CFL steps, each a minimum
    over cells of FP64 expressions that the device and the host round differently in the last
    ulp, and agrees to 1e-13 relative."""
    print(name, "device", meta, "oracle", {k: ref_meta[k] for k in ("steps", "rejected", "time", "status")}, flush=True)
    assert (meta["steps"], meta["rejected"]) == (ref_meta["steps"], ref_meta["rejected"])
    if target is None:
        assert meta["time"] == ref_meta["time"]
    else:
        assert meta["time"] < target and abs(meta["time"] - ref_meta["time"]) <= 1e-13 * ref_meta["time"]
    errs = per_variable_l2(got, ref, nv)
    print(name, "state relative L2 per variable", errs, flush=True)
    assert max(errs) <= tol


@pytest.mark.parametrize("name", list(VARIANTS))
def test_side_kind_residual_parity(name):
    case, mesh, params, u0 = variant(name)
    with Session(mesh, params, u0) as s:
        for u in (u0, perturbed(u0, mesh.n_cells)):
            errs = per_variable_l2(s.rhs(u), O.rhs(mesh, params, u), s.nv)
            print(name, "residual relative L2 per variable", errs, flush=True)
            assert max(errs) <= 1e-12


@pytest.mark.parametrize("name", list(VARIANTS))
def test_side_kind_advance_parity(name):
    case, mesh, params, u0 = variant(name)
    ref, ref_meta = O.run(mesh, params, u0, 0.01)
    assert ref_meta["status"] == 0
    with Session(mesh, params, u0) as s:
        assert s.advance(0.01) == 0
        got = s.download()
        check_run(name, got, s.meta, ref, ref_meta, s.nv)


@pytest.mark.parametrize("name", ["outlet_subsonic", "reflect_euler", "total_inlet", "noslip_side_sa"])
def test_side_kind_fp32_window(name):
    """FP32 fluxes on the same sides: within 2e-4 of the FP64 oracle at t* = 0.01."""
    case, mesh, params, u0 = variant(name)
    ref, ref_meta = O.run(mesh, params, u0, 0.01)
    with Session(mesh, params, u0, precision=32) as s:
        assert s.advance(0.01) == 0
        got = s.download()
        print(name, "fp32", s.meta, ref_meta["steps"], flush=True)
        assert abs(s.meta["steps"] - ref_meta["steps"]) <= 1 and s.meta["time"] == 0.01
    errs = per_variable_l2(got, ref, nv_of(mesh))
    print(name, "fp32 state relative L2", errs, flush=True)
    assert max(errs) <= 2e-4


@pytest.mark.parametrize("dt,tend,steps,rejected", FIXED_SCHEDULES)
def test_fixed_dt_and_rejections(dt, tend, steps, rejected):
    """Fixed steps; trials rejected for admissibility, dt halved and retried from the kept base
    residual and primitives (the oracle's counts are the compiled reference's)."""
    case, mesh, params = case_mesh("G0001", "LF1", LAMINAR)
    params.dt_fixed = dt
    u0 = mesh.initialize(params, case.freestream)
    ref, ref_meta = O.run(mesh, params, u0, tend)
    assert (ref_meta["status"], ref_meta["steps"], ref_meta["rejected"]) == (0, steps, rejected)
    with Session(mesh, params, u0) as s:
        assert s.advance(tend) == 0
        got = s.download()
        check_run(f"fixed dt {dt}", got, s.meta, ref, ref_meta, 4)


def fatal_run(mesh, params, u, tend, *, time=0.0, maxsteps=50_000_000):
    with Session(mesh, params, u, time=time) as s:
        with pytest.raises(BackendError) as e:
            s.advance(tend, maxsteps=maxsteps)
        got = s.download()
        return str(e.value), got, dict(s.meta)


def test_fatal_codes():
    case, mesh, params = case_mesh("G0001", "LF1")
    n = mesh.n_cells
    u0 = mesh.initialize(params, case.freestream)
    # 5: every one of 21 halvings of a huge fixed step fails admissibility.
    p5 = N.amr_params.from_buffer_copy(params)
    p5.dt_fixed = 1e6
    _, ref_meta = O.run(mesh, p5, u0, 1e7)
    msg, got, meta = fatal_run(mesh, p5, u0, 1e7)
    print("fatal 5", msg, meta, flush=True)
    assert "status 5" in msg and "last accepted state retained" in msg
    assert (meta["steps"], meta["rejected"]) == (ref_meta["steps"], ref_meta["rejected"]) == (0, 21)
    assert np.array_equal(got, u0)
    # 2: a nonphysical base state (negative pressure in one cell).
    bad = u0.copy()
    bad[3 * n + n // 2] = 0.1 * bad[3 * n + n // 2] - 1.0
    _, ref_meta = O.run(mesh, params, bad, 0.01)
    msg, got, meta = fatal_run(mesh, params, bad, 0.01)
    print("fatal 2", msg, meta, flush=True)
    assert "status 2" in msg and ref_meta["status"] == 1 and "nonphysical base state" in ref_meta["error"]
    assert meta["steps"] == ref_meta["steps"] == 0 and np.array_equal(got, bad)
    # 4: time + dt == time.
    _, ref_meta = O.run(mesh, params, u0, 2e20, time=1e20)
    msg, got, meta = fatal_run(mesh, params, u0, 2e20, time=1e20)
    print("fatal 4", msg, meta, flush=True)
    assert "status 4" in msg and "underflow" in ref_meta["error"]
    assert meta["steps"] == ref_meta["steps"] == 0 and meta["time"] == 1e20 and np.array_equal(got, u0)
    # 3: the step cap; the five accepted steps are kept.
    ref, ref_meta = O.run(mesh, params, u0, 0.01, maxsteps=5)
    msg, got, meta = fatal_run(mesh, params, u0, 0.01, maxsteps=5)
    print("fatal 3", msg, meta, flush=True)
    assert "status 3" in msg and "maxsteps" in ref_meta["error"]
    check_run("fatal 3", got, meta, ref, ref_meta, 5, target=0.01)


def test_session_recovers_after_a_rejected_window():
    """After a window with rejections the session advances normally from the kept state."""
    case, mesh, params = case_mesh("G0001", "LF1", LAMINAR)
    params.dt_fixed = 0.0128
    u0 = mesh.initialize(params, case.freestream)
    ref, ref_meta = O.run(mesh, params, u0, 0.1)
    ref, ref_meta = O.run(mesh, params, ref, 0.2, time=ref_meta["time"], steps=ref_meta["steps"],
                          rejected=ref_meta["rejected"])
    with Session(mesh, params, u0) as s:
        assert s.advance(0.1) == 0
        assert s.advance(0.2) == 0
        check_run("two windows", s.download(), s.meta, ref, ref_meta, 4)


def test_viscous_srd_limit():
    """SRD-grouped viscous limit on a non-agglomerated low-Reynolds laminar mesh."""
    case, mesh, params = case_mesh("G0002", "LF1", VISCOUS_SRD)
    assert mesh.info.small_cells > 0 and mesh.view.transport.stabilization == 1
    u0 = mesh.initialize(params, case.freestream)
    ref, ref_meta = O.run(mesh, params, u0, 0.01)
    assert ref_meta["status"] == 0 and ref_meta["steps"] > 33
    with Session(mesh, params, u0) as s:
        errs = per_variable_l2(s.rhs(u0), O.rhs(mesh, params, u0), 4)
        assert max(errs) <= 1e-12
        assert s.advance(0.01) == 0
        check_run("viscous srd", s.download(), s.meta, ref, ref_meta, 4)


@pytest.mark.parametrize("name", list(CONTROLS))
def test_published_control_runs(name):
    """validation/controls_checks.json:7,15,23: G0001 LF2 to t* = 0.01 in 65 / 339 / 123 steps."""
    expected = pins()["controls_t0.01"][name]
    case, mesh, params = case_mesh("G0001", "LF2", CONTROLS[name])
    u0 = mesh.initialize(params, case.freestream)
    ref, ref_meta = O.run(mesh, params, u0, 0.01)
    with Session(mesh, params, u0) as s:
        assert s.advance(0.01) == 0
        got = s.download()
        assert s.meta["steps"] == expected and s.meta["rejected"] == 0 and s.meta["time"] == 0.01
        check_run(name, got, s.meta, ref, ref_meta, s.nv)


def test_batch_isolates_failures_and_retries():
    """One arena: a fixed-dt slot with rejections, a slot that ends with fatal 5, a slot with a
    nonphysical state (fatal 2) and slots with other side kinds.  Every slot gives exactly the
    bits, counters and status of its own session."""
    items = []
    case, mesh, params = case_mesh("G0001", "LF1", LAMINAR)
    params.dt_fixed = 0.0128
    items.append(("rejections", mesh, params, mesh.initialize(params, case.freestream), 0.2))
    case, mesh, params = case_mesh("G0001", "LF1", LAMINAR)
    p5 = N.amr_params.from_buffer_copy(params)
    p5.dt_fixed = 1e6
    items.append(("fatal5", mesh, p5, mesh.initialize(params, case.freestream), 1e7))
    case, mesh, params = case_mesh("G0001", "LF1", LAMINAR)
    u = mesh.initialize(params, case.freestream)
    u[3 * mesh.n_cells + 7] = -1.0
    items.append(("fatal2", mesh, params, u, 0.2))
    for name in ("reflect_laminar", "outlet_subsonic", "noslip_side_laminar"):
        case, mesh, params, u0 = variant(name)
        items.append((name, mesh, params, u0, 0.02))
    alone = []
    for name, mesh, params, u, tend in items:
        with Session(mesh, params, u) as s:
            try:
                status = s.advance(tend)
            except BackendError:
                status = 1
            alone.append((status, s.download(), dict(s.meta)))
    with B.Batch.of_meshes([it[1] for it in items], [it[2] for it in items], [it[3] for it in items]) as b:
        b.set_targets([it[4] for it in items])
        assert b.advance("all") == 0
        states = b.states()
        for k, (name, *_rest) in enumerate(items):
            u, meta = b.download(k)
            status, ref_u, ref_meta = alone[k]
            print(name, "batch", meta, states[k]["fatal"], "alone", ref_meta, status, flush=True)
            assert meta == ref_meta
            assert (states[k]["fatal"] != 0) == (status == 1)
            assert np.array_equal(u, ref_u)
    assert [a[0] for a in alone] == [0, 1, 1, 0, 0, 0]
