"""Batch arena on the GPU: slot independence, refill, per-slot failure, steady driver.

The bar is bitwise: a mesh advanced in a slot of a batch (any neighbours, same precision and
launch tile) must give exactly the state, time and counters it gets in a session of its own,
because every kernel works on its own slot's ranges and control record and the only
cross-block reductions are exact (atomicMin of the dt bit pattern) or summed in a fixed block
order.  The single-mesh session is itself pinned to the FP64 oracle (test_gpu_parity.py);
here the FP64 batch results are also checked against the oracle directly (1e-10 relative L2).
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import batch as B
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import BackendError, Session

from .common import EULER, LAMINAR, case_mesh, merge, naca_polygon, per_variable_l2

pytestmark = pytest.mark.gpu

SRD = {"mesh": {"agglomerate_small": False}}
AUSM = {"solver": {"flux": "ausm"}}


def four_variable_meshes():
    """Three meshes of uneven size with four conserved variables (laminar, Euler SRD, laminar AUSM)."""
    out = []
    for src, preset, ov in (("G0001", "LF1", LAMINAR), ("G0010", "LF1", merge(EULER, SRD)),
                            (naca_polygon(0.12), "LF2", merge(LAMINAR, AUSM))):
        case, mesh, params = case_mesh(src, preset, ov)
        out.append((case, mesh, params, mesh.initialize(params, case.freestream)))
    return out


def sa_meshes():
    out = []
    for src, preset in (("G0002", "LF1"), ("G0100", "LF1")):
        case, mesh, params = case_mesh(src, preset)
        out.append((case, mesh, params, mesh.initialize(params, case.freestream)))
    return out


def alone(mesh, params, u0, targets, precision):
    with Session(mesh, params, u0, precision=precision) as s:
        for t in targets:
            assert s.advance(t) == 0
        return s.download(), dict(s.meta)


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_matches_alone_bitwise(precision):
    items = four_variable_meshes()
    print("cells", [m.n_cells for _, m, _, _ in items], "precision", precision, flush=True)
    refs = [alone(m, p, u, [0.01], precision) for _, m, p, u in items]
    with B.Batch.of_meshes([m for _, m, _, _ in items], [p for _, _, p, _ in items], [u for _, _, _, u in items],
                           precision=precision) as b:
        b.set_targets(0.01)
        assert b.advance("all") == 0
        for k, (ref_u, ref_meta) in enumerate(refs):
            u, meta = b.download(k)
            print(k, meta, ref_meta, flush=True)
            assert meta == ref_meta
            assert np.array_equal(u, ref_u)
        if precision == 64:
            for k, (_, m, p, u0) in enumerate(items):
                ref, ometa = O.run(m, p, u0, 0.01)
                u, meta = b.download(k)
                assert meta["steps"] == ometa["steps"]
                assert max(per_variable_l2(u, ref, 4)) <= 1e-10


def test_poll_cadence_changes_no_trajectory():
    """Polls are pipelined (the next batch of trials is queued before the host reads the last),
    so a call returns after some slots have taken extra trials; slots stop at their own
    targets, so many short calls ("any", one trial per poll) and one long call ("all", 32 per
    poll) give the same bits, times and counters, and no slot passes its target."""
    items = four_variable_meshes()
    meshes, params, states = [m for _, m, _, _ in items], [p for _, _, p, _ in items], [u for _, _, _, u in items]

    def run(until, trials):
        with B.Batch.of_meshes(meshes, params, states, precision=32, batch_trials=trials) as b:
            b.set_targets(0.01)
            calls = 0
            while True:
                running = [not (s["finished"] or s["fatal"]) for s in b.states()]
                if not any(running):
                    break
                assert b.advance(until) == 0
                calls += 1
                now = b.states()
                assert all(s["time"] <= s["target"] for s in now)
                if until == "any":  # at least one slot that was running has ended
                    assert any(r and (s["finished"] or s["fatal"]) for r, s in zip(running, now))
            return [b.download(k) for k in range(len(meshes))], calls

    ref, ref_calls = run("all", 32)
    got, calls = run("any", 1)
    print("calls", ref_calls, calls, [meta for _, meta in ref], flush=True)
    assert ref_calls == 1 and calls >= 2
    for (u, meta), (ru, rmeta) in zip(got, ref):
        assert meta == rmeta
        assert np.array_equal(u, ru)


def test_sa_batch_matches_alone_bitwise():
    items = sa_meshes()
    refs = [alone(m, p, u, [0.01], 64) for _, m, p, u in items]
    with B.Batch.of_meshes([m for _, m, _, _ in items], [p for _, _, p, _ in items],
                           [u for _, _, _, u in items]) as b:
        b.set_targets(0.01)
        assert b.advance("all") == 0
        for k, (ref_u, ref_meta) in enumerate(refs):
            u, meta = b.download(k)
            assert meta == ref_meta
            assert np.array_equal(u, ref_u)


@pytest.mark.parametrize("tile", [128, 256])
def test_fp32_sa_branches_at_production_tiles_match_single_slot_bitwise(tile):
    """The benchmark's configuration: FP32 SA-neg LF1 / LF2 meshes at tile 128 (and 256), four
    slots in three trial branches (branch 0 holds two).  Each mesh must give the bits it gets in
    a one-slot arena at the same tile (sessions use their own tile, so they are not the
    reference here)."""
    items = []
    for src, preset in (("G0002", "LF1"), ("G0100", "LF2"), ("G0001", "LF2"), ("G0010", "LF1")):
        case, mesh, params = case_mesh(src, preset)
        items.append((mesh, params, mesh.initialize(params, case.freestream)))
    refs = []
    for mesh, params, u0 in items:
        with B.Batch.of_meshes([mesh], [params], [u0], precision=32, tile=tile) as b:
            b.set_targets(0.01)
            assert b.advance("all") == 0
            refs.append(b.download(0))
    with B.Batch.of_meshes([m for m, _, _ in items], [p for _, p, _ in items], [u for _, _, u in items],
                           precision=32, tile=tile) as b:
        assert b.stats()["graph_nodes_per_trial"] > 0
        b.set_targets(0.01)
        assert b.advance("all") == 0
        for k, (ref_u, ref_meta) in enumerate(refs):
            u, meta = b.download(k)
            print(tile, k, meta, flush=True)
            assert meta == ref_meta
            assert np.array_equal(u, ref_u)


@pytest.mark.parametrize("precision", [32, 64])
def test_launch_shapes_change_no_bit(precision):
    """Flux launches with two threads per cell (the face items of a tile spread over
    more threads; each cell still sums its faces in its own order), also switched on only once
    few slots run and without graphs: every slot ends with the bits of the default shape."""
    items = []
    for src, preset in (("G0002", "LF1"), ("G0100", "LF2"), ("G0001", "LF2"), ("G0010", "LF1")):
        case, mesh, params = case_mesh(src, preset)
        items.append((mesh, params, mesh.initialize(params, case.freestream)))
    targets = [0.004, 0.006, 0.01, 0.01]

    def run(shape, graph=True, tile=128):
        with B.Batch.of_meshes([m for m, _, _ in items], [p for _, p, _ in items], [u for _, _, u in items],
                               precision=precision, tile=tile, graph=graph) as b:
            if shape is not None:
                b.set_launch_shape(*shape)
            b.set_targets(targets)
            assert b.advance("all") == 0
            return [b.download(k) for k in range(len(items))]

    refs = {}  # the default shape at each tile (another tile rounds the face sums in another order)
    for shape, graph, tile in (((2, 0, 2), True, 128), ((1, 2, 2), True, 128), ((1, 3, 2), False, 128),
                               ((2, 0, 2), True, 64)):
        print(precision, "shape", shape, "graph", graph, "tile", tile, flush=True)
        if tile not in refs:
            refs[tile] = run(None, True, tile)
        for k, ((u, meta), (ref_u, ref_meta)) in enumerate(zip(run(shape, graph, tile), refs[tile])):
            assert meta == ref_meta, (shape, k)
            assert np.array_equal(u, ref_u), (shape, k)
    with B.Batch.of_meshes([items[0][0]], [items[0][1]], [items[0][2]], precision=precision, tile=128) as b:
        with pytest.raises(BackendError):
            b.set_launch_shape(1, 4, 4)  # four threads per cell


def test_refill_and_release_keep_neighbours_bitwise():
    items = four_variable_meshes()
    (ca, ma, pa, ua), (cb, mb, pb, ub), (cc, mc, pc, uc) = items
    caps, feats = zip(*[B.capacity(m) for _, m, _, _ in items])
    cap = B.envelope(caps)
    ref_a = alone(ma, pa, ua, [0.005, 0.01], 64)
    ref_c = alone(mc, pc, uc, [0.01], 64)
    with B.Batch.with_slots([cap, cap, cap], B.union(feats)) as b:
        assert b.stats()["allocated_bytes"] == B.arena_bytes([cap] * 3, B.union(feats))
        b.admit(0, ma, pa, ua)
        b.admit(1, mb, pb, ub)
        b.set_targets([0.005, 0.005, 0.005])
        assert b.advance("all") == 0
        b.release(1)
        assert b.states()[1]["n_cells"] == 0
        b.admit(2, mc, pc, uc)
        b.set_targets(0.01)
        assert b.advance("all") == 0
        u, meta = b.download(0)
        assert meta == ref_a[1] and np.array_equal(u, ref_a[0])
        u, meta = b.download(2)
        assert meta == ref_c[1] and np.array_equal(u, ref_c[0])
        st = b.states()
        assert st[1]["n_cells"] == 0 and st[1]["steps"] == 0


def test_oversized_admission_leaves_slot_untouched():
    items = four_variable_meshes()
    (_, ma, pa, ua), _, (_, mc, pc, uc) = items
    cap, feat = B.capacity(ma)
    _, feat_c = B.capacity(mc)
    with B.Batch.with_slots([cap], B.union([feat, feat_c])) as b:
        b.admit(0, ma, pa, ua)
        b.set_targets(0.002)
        b.advance("all")
        before, meta = b.download(0)
        with pytest.raises(BackendError, match="capacity"):
            b.admit(0, mc, pc, uc)
        after, meta2 = b.download(0)
        assert meta == meta2 and np.array_equal(before, after)


def test_failed_slot_does_not_disturb_others():
    items = four_variable_meshes()
    ref = alone(items[1][1], items[1][2], items[1][3], [0.01], 64)
    with B.Batch.of_meshes([m for _, m, _, _ in items[:2]], [p for _, _, p, _ in items[:2]],
                           [u for _, _, _, u in items[:2]]) as b:
        b.set_targets(0.01, maxsteps=[5, 10 ** 9])
        assert b.advance("all") == 0
        st = b.states()
        assert st[0]["fatal"] == 3 and st[0]["steps"] == 5
        assert st[1]["fatal"] == 0 and st[1]["finished"]
        u, meta = b.download(1)
        assert meta == ref[1] and np.array_equal(u, ref[0])


def test_monitor_matches_session_bitwise():
    items = four_variable_meshes()
    frame = [-1.0, 3.0, -1.0, 1.0, 0.25, 0.0]
    with B.Batch.of_meshes([m for _, m, _, _ in items], [p for _, _, p, _ in items],
                           [u for _, _, _, u in items]) as b:
        b.set_targets(0.01)
        b.advance("all")
        out, status = b.monitor([1, 0, 1], [frame] * 3, [c.mach for c, _, _, _ in items],
                                [c.aoa for c, _, _, _ in items])
        assert list(status) == [0, 3, 0]
        for k in (0, 2):
            case, mesh, params, _ = items[k]
            u, meta = b.download(k)
            with Session(mesh, params, u, time=meta["time"], steps=meta["steps"]) as s:
                ref = s.monitor(tuple(frame[:4]), tuple(frame[4:]), case.mach, case.aoa)
            assert np.array_equal(out[k, :5], ref["residual"])
            assert (out[k, 5], out[k, 6], out[k, 7]) == (ref["cl"], ref["cd"], ref["cm"])


def test_driver_matches_session_solve_bitwise():
    items = four_variable_meshes()[:2]
    refs = []
    for case, mesh, params, u0 in items:
        with Session(mesh, params, u0) as s:
            refs.append(s.solve(case, mode="time", stop_time=0.05, budget=60, history=0))
    with B.Batch.of_meshes([m for _, m, _, _ in items], [p for _, _, p, _ in items],
                           [u for _, _, _, u in items]) as b:
        results = b.solve([c for c, _, _, _ in items], mode="time", budget=60, stop_time=0.05)
    for r, ref in zip(results, refs):
        print(r["steps"], r["cl"], ref["steps"], ref["cl"], flush=True)
        for key in ("steps", "rejected", "time", "cl", "cd", "cm", "residual", "checks", "reason"):
            assert r[key] == ref[key], key


@pytest.mark.parametrize("tile", [128, 256])
def test_fp64_batch_at_large_tiles_matches_oracle(tile):
    """The larger launch tiles (whose face-item chunks and XCD runs differ from tile 64's)
    against the FP64 oracle over t* = 0.01 on an SA-neg and a laminar AUSM mesh."""
    items = sa_meshes()[:1] + four_variable_meshes()[2:]
    # One arena per conserved-variable count: SA-neg (5) and laminar (4).
    for case, mesh, params, u0 in items:
        with B.Batch.of_meshes([mesh], [params], [u0], precision=64, tile=tile) as b:
            b.set_targets(0.01)
            assert b.advance("all") == 0
            u, meta = b.download(0)
        ref, ometa = O.run(mesh, params, u0, 0.01)
        nv = 5 if case.sa_model else 4
        err = max(per_variable_l2(u, ref, nv))
        print("tile", tile, "cells", mesh.n_cells, "steps", meta["steps"], ometa["steps"], "rel L2", err, flush=True)
        assert meta["steps"] == ometa["steps"]
        assert err <= 1e-10
