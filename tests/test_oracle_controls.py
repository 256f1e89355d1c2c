"""Pins of the oracle's boundary states and step-control branches (oracle/amr_oracle.cpp).

1. Boundary point functions (`boundary_state`, `boundary_valid`) bit for bit against the
   reference's own amr_bc_state / amr_bc_valid (src/static_amr2d_boundary.h:32-102, compiled from
   its sources into oracle/_ref) over random interior states, every kind, every side.
2. Closed forms of the same functions: transmissive = interior, reflect mirrors the normal
   velocity, no-slip mirrors about the wall velocity, supersonic inflow = the reference state,
   supersonic outflow = the interior state, subsonic characteristic states keep the outgoing
   Riemann invariant, the incoming one (farfield) or the outlet pressure (pressure outlet) and the
   upwind entropy, and a total inlet's ghost state has the prescribed total pressure, total
   temperature and direction.
3. Whole residuals and runs on meshes with every domain-side kind (pressure outlet, reflect,
   transmissive, total inlet, no-slip side, inherited legacy kind) bit for bit against the
   compiled reference.
4. Step control bit for bit against the compiled reference: fixed dt, rejected trials with dt
   halving, 21 rejections (fatal 5), a nonphysical base state (fatal 2), dt underflow (fatal 4),
   the step cap (fatal 3), and the SRD-mapped viscous limit of a non-agglomerated mesh.
5. The reference's published control runs (validation/controls_checks.json:7,15,23): G0001 LF2
   to t* = 0.01 in 65 steps (subsonic pressure outlet), 339 (fixed dt 1e-7 s) and 123 (AUSM+-up,
   CFL 0.1), no rejection.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import _native as N

from .common import LAMINAR, case_mesh, pins
from .variants import CONTROLS, FIXED_SCHEDULES, GAMMA, GM, VARIANTS, VISCOUS_SRD

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference CPU build (oracle/_ref) not present")

def both(name, sides=None):
    src, preset, over, sd = VARIANTS[name] if isinstance(name, str) else name
    case, mine, params = case_mesh(src, preset, over, sides=sd)
    _, ref, _ = case_mesh(src, preset, over, lib=O.ref_lib(), prefix="ref_mesh", sides=sd)
    return case, mine, ref, params


def perturbed(u, n, seed=5):
    rng = np.random.default_rng(seed)
    v = u.copy()
    v[:n] *= 1 + 0.02 * rng.standard_normal(n)
    v[n:2 * n] += 0.05 * rng.standard_normal(n)
    v[2 * n:3 * n] += 0.05 * rng.standard_normal(n)
    return v


def same_run(a, ma, b, mb):
    assert (ma["status"], ma["error"]) == (mb["status"], mb["error"])
    assert (ma["steps"], ma["rejected"], ma["time"]) == (mb["steps"], mb["rejected"], mb["time"])
    assert np.array_equal(a, b)


# ------------------------------------------------------------------ 1. point functions ----
def random_boundary(rng, kind):
    b = N.amr_boundary()
    b.kind = kind
    ref = (rng.uniform(0.5, 2.0), rng.uniform(-2.0, 2.0), rng.uniform(-1.0, 1.0), rng.uniform(0.3, 1.5))
    for v in range(4):
        b.reference[v] = ref[v]
    b.total_temperature = rng.uniform(0.6, 1.5)
    b.total_pressure = rng.uniform(0.8, 3.0)
    b.gas_constant = rng.uniform(0.8, 1.2)
    ang = rng.uniform(-0.6, 0.6)
    b.direction[0], b.direction[1] = math.cos(ang), math.sin(ang)
    b.outlet_pressure = rng.uniform(0.3, 1.5)
    b.wall_velocity[0], b.wall_velocity[1] = rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5)
    return b


def random_interior(rng):
    r, p = rng.uniform(0.3, 2.0), rng.uniform(0.2, 1.5)
    a = math.sqrt(GAMMA * p / r)
    # Normal Mach numbers from -2.5 to 2.5 cover supersonic inflow, subsonic and supersonic outflow.
    return np.array([r, rng.uniform(-2.5, 2.5) * a, rng.uniform(-2.5, 2.5) * a, p])


@needs_ref
@pytest.mark.parametrize("kind", [-1, 1, 2, 3, 4, 5, 6, 7, -2])
def test_boundary_state_bit_identical_to_reference(kind):
    rng = np.random.default_rng(100 + kind)
    outcomes = set()
    for trial in range(400):
        b = random_boundary(rng, kind)
        side = trial % 4
        legacy = int(rng.integers(1, 7))
        if trial % 7 == 0:  # valid direction / wall velocity for this side
            b.direction[side // 2] = -abs(b.direction[side // 2]) if side % 2 else abs(b.direction[side // 2])
            b.wall_velocity[side // 2] = 0.0
        wi = random_interior(rng)
        if trial % 11 == 0:
            wi[3] = -wi[3]  # inadmissible interior state
        ok, wb = O.boundary_state(b, legacy, GAMMA, side, wi)
        rok, rwb = O.ref_boundary_state(b, legacy, GAMMA, side, wi)
        assert ok == rok
        assert np.array_equal(wb, rwb)
        assert O.boundary_valid(b, legacy, side) == O.ref_boundary_valid(b, legacy, side)
        outcomes.add(ok)
    if kind in (1, 2, 3, 5, 6):
        assert True in outcomes


# ------------------------------------------------------------------ 2. closed forms -------
def frame(side, w):
    """Outward normal velocity, tangential velocity, sound speed of a state on `side`."""
    axis = side // 2
    sign = 1.0 if side % 2 else -1.0
    return sign * w[1 + axis], w[2 - axis], math.sqrt(GAMMA * w[3] / w[0])


def boundary(kind, **fields):
    b = N.amr_boundary()
    b.kind = kind
    for v, x in enumerate(fields.get("reference", (1.0, 0.5, 0.1, 1 / GAMMA))):
        b.reference[v] = x
    b.total_pressure, b.total_temperature, b.gas_constant = fields.get("total", (0.0, 0.0, 1.0))
    b.direction[0], b.direction[1] = fields.get("direction", (1.0, 0.0))
    b.outlet_pressure = fields.get("outlet", 0.0)
    b.wall_velocity[0], b.wall_velocity[1] = fields.get("wall", (0.0, 0.0))
    return b


@pytest.mark.parametrize("side", range(4))
def test_boundary_wall_and_copy_kinds(side):
    wi = np.array([0.9, 0.4, -0.3, 0.8])
    ok, wb = O.boundary_state(boundary(1), 1, GAMMA, side, wi)
    assert ok and np.array_equal(wb, wi)  # transmissive: zero-gradient copy
    ok, wb = O.boundary_state(boundary(2), 1, GAMMA, side, wi)
    mirrored = wi.copy()
    mirrored[1 + side // 2] = -wi[1 + side // 2]
    assert ok and np.array_equal(wb, mirrored)  # reflect: normal velocity mirrored
    ok, wb = O.boundary_state(boundary(6, wall=(0.25, -0.125)), 1, GAMMA, side, wi)
    assert ok and np.array_equal(wb, [wi[0], 0.5 - wi[1], -0.25 - wi[2], wi[3]])  # mean velocity = wall
    ok, wb = O.boundary_state(boundary(-1), 2, GAMMA, side, wi)  # inherited legacy kind
    assert ok and np.array_equal(wb, mirrored)


@pytest.mark.parametrize("kind", [3, 5])
@pytest.mark.parametrize("side", range(4))
def test_boundary_supersonic_in_and_outflow(kind, side):
    ref = (1.2, 0.3, -0.2, 0.9)
    b = boundary(kind, reference=ref, outlet=0.7)
    sign = 1.0 if side % 2 else -1.0
    w_in = np.array([1.0, 0.1, 0.1, 0.7])
    w_in[1 + side // 2] = -sign * 3.0  # |un| = 3 > a = 0.99: supersonic inflow
    ok, wb = O.boundary_state(b, 1, GAMMA, side, w_in)
    assert ok and np.array_equal(wb, ref)
    w_out = w_in.copy()
    w_out[1 + side // 2] = sign * 3.0  # supersonic outflow: every characteristic leaves
    ok, wb = O.boundary_state(b, 1, GAMMA, side, w_out)
    assert ok and np.array_equal(wb, w_out)


@pytest.mark.parametrize("side", range(4))
@pytest.mark.parametrize("mach_n", [-0.6, -0.1, 0.2, 0.7])
def test_farfield_subsonic_characteristics(side, mach_n):
    ref = (1.0, 0.4, 0.2, 1 / GAMMA)
    wi = np.array([0.9, 0.35, 0.25, 0.65])
    ai = math.sqrt(GAMMA * wi[3] / wi[0])
    sign = 1.0 if side % 2 else -1.0
    wi[1 + side // 2] = sign * mach_n * ai
    ok, wb = O.boundary_state(boundary(3, reference=ref), 1, GAMMA, side, wi)
    assert ok
    un_i, _, a_i = frame(side, wi)
    un_r, _, a_r = frame(side, np.array(ref))
    un_b, ut_b, a_b = frame(side, wb)
    assert un_b + 2 * a_b / GM == pytest.approx(un_i + 2 * a_i / GM, rel=1e-13)  # outgoing invariant
    assert un_b - 2 * a_b / GM == pytest.approx(un_r - 2 * a_r / GM, rel=1e-13)  # incoming invariant
    up = wi if un_i >= 0 else np.array(ref)  # entropy and tangential velocity from upwind
    assert wb[3] / wb[0] ** GAMMA == pytest.approx(up[3] / up[0] ** GAMMA, rel=1e-13)
    assert ut_b == up[2 - side // 2]


@pytest.mark.parametrize("side", range(4))
def test_pressure_outlet_subsonic_outflow(side):
    p_out = 0.6
    wi = np.array([1.1, 0.2, 0.1, 0.75])
    sign = 1.0 if side % 2 else -1.0
    wi[1 + side // 2] = sign * 0.4
    ok, wb = O.boundary_state(boundary(5, outlet=p_out), 1, GAMMA, side, wi)
    assert ok
    un_i, ut_i, a_i = frame(side, wi)
    un_b, ut_b, a_b = frame(side, wb)
    assert wb[3] == p_out  # the outlet pressure is imposed
    assert wb[3] / wb[0] ** GAMMA == pytest.approx(wi[3] / wi[0] ** GAMMA, rel=1e-13)  # isentropic
    assert un_b + 2 * a_b / GM == pytest.approx(un_i + 2 * a_i / GM, rel=1e-13)  # outgoing invariant
    assert ut_b == ut_i
    # Reversed flow at the outlet is a farfield state with the reference as the reservoir.
    wi[1 + side // 2] = -sign * 0.3
    ok, wb = O.boundary_state(boundary(5, outlet=p_out), 1, GAMMA, side, wi)
    ok3, wb3 = O.boundary_state(boundary(3), 1, GAMMA, side, wi)
    assert ok and ok3 and np.array_equal(wb, wb3)


@pytest.mark.parametrize("side", range(4))
def test_total_inlet_total_conditions(side):
    pt, tt, gas = 1.2, 0.9, 1.0
    axis = side // 2
    sign = 1.0 if side % 2 else -1.0
    d = np.zeros(2)
    d[axis], d[1 - axis] = -sign * math.cos(0.3), math.sin(0.3)  # pointing into the domain
    b = boundary(4, total=(pt, tt, gas), direction=tuple(d))
    assert O.boundary_valid(b, 1, side)
    wi = np.array([1.0, 0.0, 0.0, 0.8])
    wi[1 + axis] = -sign * 0.3  # subsonic inflow
    ok, wb = O.boundary_state(b, 1, GAMMA, side, wi)
    assert ok
    temp = wb[3] / (wb[0] * gas)
    speed2 = wb[1] ** 2 + wb[2] ** 2
    assert temp + GM / (2 * GAMMA * gas) * speed2 == pytest.approx(tt, rel=1e-14)  # total temperature
    assert wb[3] * (tt / temp) ** (GAMMA / GM) == pytest.approx(pt, rel=1e-13)  # total pressure
    v = math.sqrt(speed2)
    assert wb[1] == pytest.approx(v * d[0], rel=1e-15, abs=1e-300) and wb[2] == pytest.approx(v * d[1], rel=1e-15)
    un_i, _, a_i = frame(side, wi)
    un_b, _, a_b = frame(side, wb)
    assert un_b + 2 * a_b / GM == pytest.approx(un_i + 2 * a_i / GM, rel=1e-13)  # outgoing invariant
    wi[1 + axis] = sign * 0.3  # outflow through an inlet is a failure, not a clipped state
    assert not O.boundary_state(b, 1, GAMMA, side, wi)[0]


# ------------------------------------------------------------------ 3. mesh variants ------
@needs_ref
@pytest.mark.parametrize("name", list(VARIANTS))
def test_rhs_and_run_bit_identical_with_side_kinds(name):
    case, mine, ref, params = both(name)
    n = mine.n_cells
    u0 = mine.initialize(params, case.freestream)
    for u in (u0, perturbed(u0, n)):
        assert np.array_equal(O.rhs(mine, params, u), O.ref_rhs(ref, params, u))
    a, ma = O.run(mine, params, u0, 0.01)
    b, mb = O.ref_run(ref, params, u0, 0.01)
    assert ma["status"] == 0
    same_run(a, ma, b, mb)


@needs_ref
def test_noslip_side_changes_the_solution():
    """The side kind reaches the residual: a no-slip side differs from the farfield default."""
    case, mesh, params = case_mesh("G0001", "LF1", LAMINAR, sides=VARIANTS["noslip_side_laminar"][3])
    _, plain, _ = case_mesh("G0001", "LF1", LAMINAR)
    u = perturbed(mesh.initialize(params, case.freestream), mesh.n_cells)
    assert not np.array_equal(O.rhs(mesh, params, u), O.rhs(plain, params, u))


# ------------------------------------------------------------------ 4. step control -------
@needs_ref
@pytest.mark.parametrize("dt,tend,steps,rejected", FIXED_SCHEDULES)
def test_fixed_dt_and_rejections_bit_identical(dt, tend, steps, rejected):
    case, mine, ref, params = both(("G0001", "LF1", LAMINAR, None))
    params.dt_fixed = dt
    u0 = mine.initialize(params, case.freestream)
    a, ma = O.run(mine, params, u0, tend)
    b, mb = O.ref_run(ref, params, u0, tend)
    same_run(a, ma, b, mb)
    assert (ma["status"], ma["steps"], ma["rejected"], ma["time"]) == (0, steps, rejected, tend)


@needs_ref
def test_fatal_codes_bit_identical():
    case, mine, ref, params = both(("G0001", "LF1", None, None))
    n = mine.n_cells
    u0 = mine.initialize(params, case.freestream)
    # 5: every one of 21 halvings of a huge fixed step fails admissibility.
    p5 = N.amr_params.from_buffer_copy(params)
    p5.dt_fixed = 1e6
    a, ma = O.run(mine, p5, u0, 1e7)
    b, mb = O.ref_run(ref, p5, u0, 1e7)
    same_run(a, ma, b, mb)
    assert ma["status"] == 1 and ma["rejected"] == 21 and ma["steps"] == 0 and "21 trials" in ma["error"]
    assert np.array_equal(a, u0)  # the last accepted state is kept
    # 2: a nonphysical base state (negative pressure in one cell).
    bad = u0.copy()
    bad[3 * n + n // 2] = 0.1 * bad[3 * n + n // 2] - 1.0
    a, ma = O.run(mine, params, bad, 0.01)
    b, mb = O.ref_run(ref, params, bad, 0.01)
    same_run(a, ma, b, mb)
    assert ma["status"] == 1 and ma["steps"] == 0 and "nonphysical base state" in ma["error"]
    # 4: time + dt == time.
    a, ma = O.run(mine, params, u0, 2e20, time=1e20)
    b, mb = O.ref_run(ref, params, u0, 2e20, time=1e20)
    same_run(a, ma, b, mb)
    assert ma["status"] == 1 and ma["steps"] == 0 and "underflow" in ma["error"]
    # 3: the step cap.
    a, ma = O.run(mine, params, u0, 0.01, maxsteps=5)
    b, mb = O.ref_run(ref, params, u0, 0.01, maxsteps=5)
    same_run(a, ma, b, mb)
    assert ma["status"] == 1 and ma["steps"] == 5 and "maxsteps" in ma["error"]


@needs_ref
def test_viscous_srd_limit_bit_identical():
    case, mine, ref, params = both(("G0002", "LF1", VISCOUS_SRD, None))
    assert mine.info.small_cells > 0 and mine.view.transport.stabilization == 1
    u0 = mine.initialize(params, case.freestream)
    a, ma = O.run(mine, params, u0, 0.01)
    b, mb = O.ref_run(ref, params, u0, 0.01)
    same_run(a, ma, b, mb)
    # 33 steps when the convective limit binds (mu = 1e-3 Pa s); the SRD-mapped viscous limit
    # takes more, the true-volume limit of the smallest cut cells far more (5058).
    assert ma["status"] == 0 and ma["steps"] > 33


def test_sa_with_srd_is_rejected_by_the_mesher():
    """SA-neg needs an agglomerated true-volume mesh (src/static_amr2d_sa.cpp:8)."""
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.mesh import MeshError
    with pytest.raises(MeshError, match="SA-neg requires agglomerated"):
        case_mesh("G0001", "LF1", {"physics": {"viscous_stabilization": "srd"}})


# ------------------------------------------------------------------ 5. published controls --
@pytest.mark.parametrize("name", list(CONTROLS))
def test_published_control_runs(name):
    expected = pins()["controls_t0.01"][name]
    case, mesh, params = case_mesh("G0001", "LF2", CONTROLS[name])
    _, meta = O.run(mesh, params, mesh.initialize(params, case.freestream), 0.01)
    assert meta["status"] == 0 and meta["rejected"] == 0
    assert meta["steps"] == expected and meta["time"] == 0.01
