"""The ragged batch and the serving loop at test scale on the GPU.

Ragged batch: airfoils drawn as the workload draws them (thickness U[0.06, 0.14],
camber U[0, 0.04], camber position U[0.3, 0.5], M U[1.2, 2.0], AoA U[-2, 4] deg), meshes LF1 or
LF2 50/50, SA-neg, FP32 work arrays, tile 128, three trial branches.  Every slot equals a
one-slot arena bit for bit and the FP64 oracle within 2e-4 relative L2 at t* = 0.01 (step counts
within 1); a 256-mesh mixed arena is advanced a fixed number of trials and every slot checked.

Serving: SA-neg requests on LF1 / LF2 / level-7 meshes served to steady from an arena
of three slot classes sized through the HBM-budget path; each result equals the same request
solved alone, bit for bit.  An FP64 serving window matches the oracle within 1e-10.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import pyoracle as O
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import batch as B
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import config as CF
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import geometry as G
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import mesh as M
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import serve as S
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import Session

from .common import per_variable_l2

pytestmark = pytest.mark.gpu


def draw(n, seed):
    rng = np.random.default_rng(seed)
    return [dict(thickness=float(rng.uniform(0.06, 0.14)), camber=float(rng.uniform(0.0, 0.04)),
                 position=float(rng.uniform(0.3, 0.5)), mach=float(rng.uniform(1.2, 2.0)),
                 aoa=float(rng.uniform(-2.0, 4.0))) for _ in range(n)]


def polygon(spec):
    return G.reorder_points(G.naca4_points(spec["thickness"], spec["camber"], spec["position"], 100))


def case_of(spec, preset):
    return CF.make_case(preset, {"freestream": {"Mach": spec["mach"], "AOA_degrees": spec["aoa"]}})


def build(spec, preset):
    case = case_of(spec, preset)
    poly, patches = polygon(spec)
    mesh = M.build_case_mesh(case, poly, patches)
    params = M.params_of(case)
    return case, mesh, params, mesh.initialize(params, case.freestream)


def mixed(n, seed):
    specs = draw(n, seed)
    presets = ["LF1" if p < 0.5 else "LF2" for p in np.random.default_rng(seed + 1).random(n)]
    with ThreadPoolExecutor(16) as pool:
        return list(pool.map(build, specs, presets)), presets


def one_slot(mesh, params, u0, tile, target=None, trials=None):
    with B.Batch.of_meshes([mesh], [params], [u0], precision=32, tile=tile) as b:
        if trials is None:
            b.set_targets(target)
            assert b.advance("all") == 0
        else:
            b.set_targets(1e30)
            b.run(trials)
        return b.download(0)


def test_ragged_mixed_batch_fp32():
    items, presets = mixed(16, 303)
    print("ragged batch: presets", presets, "cells", [m.n_cells for _, m, _, _ in items], flush=True)
    assert 4 <= presets.count("LF1") <= 12
    with B.Batch.of_meshes([m for _, m, _, _ in items], [p for _, _, p, _ in items], [u for _, _, _, u in items],
                           precision=32, tile=128) as b:
        b.set_targets(0.01)
        assert b.advance("all") == 0
        got = [b.download(k) for k in range(len(items))]
    worst = 0.0
    for k, ((case, mesh, params, u0), (u, meta)) in enumerate(zip(items, got)):
        ref_u, ref_meta = one_slot(mesh, params, u0, 128, target=0.01)
        assert meta == ref_meta and np.array_equal(u, ref_u), k
        oracle_u, oracle_meta = O.run(mesh, params, u0, 0.01)
        assert oracle_meta["status"] == 0 and meta["time"] == 0.01
        assert abs(meta["steps"] - oracle_meta["steps"]) <= 1, (k, meta, oracle_meta)
        err = max(per_variable_l2(u, oracle_u, 5))
        worst = max(worst, err)
        assert err <= 2e-4, (k, err)
    print("ragged batch: worst FP32 vs FP64 oracle relative L2", worst, flush=True)


def test_ragged_256_mesh_arena_fixed_trials():
    items, presets = mixed(256, 2027)
    trials = 120
    print("ragged batch: 256 meshes,", presets.count("LF1"), "LF1,", sum(m.n_cells for _, m, _, _ in items), "cells",
          flush=True)
    with B.Batch.of_meshes([m for _, m, _, _ in items], [p for _, _, p, _ in items], [u for _, _, _, u in items],
                           precision=32, tile=128) as b:
        b.set_targets(1e30)
        states = b.run(trials)
        picks = [0, 97, 255]
        got = {k: b.download(k) for k in picks}
        for k, st in enumerate(states):
            u, meta = b.download(k)
            nv, n = 5, items[k][1].n_cells
            q = u.reshape(nv, n)
            p = 0.4 * (q[3] - 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0])
            assert not st["fatal"] and st["steps"] + st["rejected"] == trials and st["rejected"] == 0
            assert 0 < st["time"] < 1e30 and np.all(np.isfinite(u)) and q[0].min() > 0 and p.min() > 0, k
    for k in picks:
        case, mesh, params, u0 = items[k]
        ref_u, ref_meta = one_slot(mesh, params, u0, 128, trials=trials)
        assert got[k][1] == ref_meta and np.array_equal(got[k][0], ref_u), k


SERVE_MIX = (("LF1", 2), ("LF2", 2), ("L7", 1))


def test_serving_sa_neg_to_steady_equals_alone():
    specs = draw(5, 505)
    presets = [p for p, count in SERVE_MIX for _ in range(count)]
    arrivals = S.poisson_arrivals(5.0, len(specs), seed=11)
    reqs = [S.Request(i, polygon(specs[i]), case_of(specs[i], presets[i]), float(arrivals[i])) for i in range(5)]
    built = [S.build_request_mesh(r, cache=None) for r in reqs]
    # Three slot classes: the capacity of each preset's meshes (+5 %), as many slots of each as
    # an HBM budget of three level-7 slots allows, at most two.
    caps, feats = {}, []
    for preset, _ in SERVE_MIX:
        cs, fs = zip(*[B.capacity(m, 64) for (m, _, _), p in zip(built, presets) if p == preset])
        caps[preset] = B.envelope(cs, margin=0.05, tile=64)
        feats += fs
    features = B.union(feats)
    budget = 3 * B.arena_bytes([caps["L7"]], features, 32, 64)
    slot_caps = []
    for preset, _ in SERVE_MIX:
        fit = B.slots_fitting(caps[preset], features, budget // 3, precision=32, tile=64)
        slot_caps += [caps[preset]] * min(2, fit)
    assert len(slot_caps) >= 3 and B.arena_bytes(slot_caps, features, 32, 64) <= budget * 1.01
    print("serving: slots", len(slot_caps), "cells", [m.n_cells for m, _, _ in built], flush=True)
    backend = S.DeviceBackend(slot_caps, features, precision=32, tile=64)
    try:
        rep = S.Server(backend, mode="steady", budget=600, workers=4,
                       mesher=lambda r: S.build_request_mesh(r, cache=None)).serve(reqs)
    finally:
        backend.close()
    print({k: v for k, v in rep.items() if k != "outcomes"}, flush=True)
    assert rep["completed"] == 5 and rep["converged"] == 5
    for req, (mesh, params, u0), out in zip(reqs, built, rep["outcomes"]):
        with Session(mesh, params, u0, precision=32) as s:
            ref = s.solve(req.case, mode="steady", budget=600, history=0)
        got = out.result
        print(req.rid, presets[req.rid], out.slot, got["steps"], ref["steps"], got["cl"], ref["cl"], out.latency,
              flush=True)
        assert got["converged"] and got["stop_reason"] == ref["stop_reason"]
        for key in ("steps", "rejected", "time", "cl", "cd", "cm", "residual", "checks"):
            assert got[key] == ref[key], key


def test_fp64_serving_window_matches_oracle():
    specs = draw(3, 55)
    presets = ["LF1", "LF2", "LF1"]
    reqs = [S.Request(i, polygon(specs[i]), case_of(specs[i], presets[i]), 0.01 * i) for i in range(3)]
    built = [S.build_request_mesh(r, cache=None) for r in reqs]
    caps, feats = zip(*[B.capacity(m) for m, _, _ in built])
    backend = S.DeviceBackend([B.envelope(caps)] * 2, B.union(feats), precision=64)
    finals = {}
    try:
        server = S.Server(backend, mode="time", stop_time=0.01, budget=120, workers=2,
                          mesher=lambda r: S.build_request_mesh(r, cache=None))
        rep = server.serve(reqs)
    finally:
        backend.close()
    assert rep["completed"] == 3
    for req, (mesh, params, u0), out in zip(reqs, built, rep["outcomes"]):
        ref, meta = O.run(mesh, params, u0, 0.01)
        conv = req.case.convergence
        mon = O.monitor(mesh, params, ref, conv["residual_region_xy_c"], req.case.moment_reference, req.case.mach,
                        req.case.aoa)
        got = out.result
        print(req.rid, got["steps"], meta["steps"], got["cl"], mon["cl"], flush=True)
        assert got["steps"] == meta["steps"] and got["time"] == 0.01 and got["checks"] == 1
        assert np.allclose(got["residual"][:5], mon["residual"][:5], rtol=1e-10, atol=0)
        for key in ("cl", "cd", "cm"):
            assert abs(got[key] - mon[key]) <= 1e-10, key
        finals[req.rid] = got
    assert len(finals) == 3
