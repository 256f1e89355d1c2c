#!/usr/bin/env python3
"""Benchmark: batched FP32 LF2 SA-neg steady airfoil solves per hour on one MI355X.

Workload (synthetic, seeded): --batch airfoils (default 256) drawn from
numpy.random.default_rng(--seed): NACA-4-style thickness U[0.06, 0.14], camber U[0, 0.04],
camber position U[0.3, 0.5], Mach U[1.2, 2.0], AoA U[-2, 4] deg; LF2 meshes (6 levels, wall
h/c 1/128, ~9.5k cells each); SA-neg RANS with Sutherland viscosity, HLLC, CFL 0.15
(viscous/SA 0.1); uniform freestream start.  All meshes live in one device arena and advance
together, one hipGraph replay per step; each mesh stops at the steady criteria of the
reference (raw residual <= 1e-3, coefficient changes <= 0.005 for 5 checks, t* >= 3).

A "step" is one SSP-RK3 step of the whole batch (one trial of every mesh).  --warmup untimed
steps, then exactly --steps steps timed between two device synchronisations: ms_per_step.

value = batch solves per hour measured directly: 3600 * (meshes that reached steady) / (wall
seconds of the batched steady solve of the whole batch, from the freestream start, polling
and on-device checks included).  The plain path is measured the same way: one FP64 session
(the drop-in of static_amr2d_gpu.h) on the single LF2 case (thickness 10%, camber 2% at
0.4c, M 1.65, AoA 1.5 deg) to steady.  Every batched FP32 result is compared with the same
mesh solved in FP64 (every mesh by default, --fp64-check; an FP64 batch, bitwise equal to FP64
sessions, untimed): max and p99 of |dCL|, |dCD|, |dCM|, and the rejected trials of both runs.

latency: a bounded serving leg (--serve-leg requests, seeded airfoils on LF1 / LF2 / level-7
meshes, 50/35/15 %, Poisson arrivals at 0.8 x the estimated capacity, continuous batching):
p50 / p99 time-to-steady and its solves/hour.

roofline: the stage pass (gradient + fused flux/gather/RK kernel of the last RK stage) of the whole batch,
per-kernel hipEvent times on the arena's stream; algorithmic bytes are the contract bytes of
roofline.py (single-width state), with this build's own byte count (FP64 state) beside them.
cpu_baseline: the FP64 oracle on the host cores (one thread, the reference's default build, and
OMP_NUM_THREADS threads with the reference's OpenMP loops), a bounded sample of steps of the
single LF2 case, scaled by that case's FP64 steps to steady.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "LF2 SA-neg steady solves/hour batched FP32 vs plain FP64; p50/p99 latency; HBM GB/s"
PKG = "quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


PROFILES = Path(__file__).resolve().parent / "profiles"


def profiled_traffic(batch: int, tile: int, seed: int):
    """HBM bytes of one stage pass (gradient_stage + flux_rk2) of this workload as measured by
    rocprofv3 (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a copy kernel of known size),
    read from the committed profile summary; None when the summary covers another workload."""
    for path in sorted(PROFILES.glob("stage_pass_*.md")):
        text = path.read_text()
        w = re.search(r"^stage_pass_profile: batch=(\d+) tile=(\d+) seed=(\d+)$", text, re.M)
        m = re.search(r"^stage_pass_hbm_bytes: (\d+)$", text, re.M)
        if w and m and tuple(int(x) for x in w.groups()) == (batch, tile, seed):
            return int(m.group(1))
    return None


def draw_airfoils(n: int, seed: int) -> list[dict]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append(dict(thickness=float(rng.uniform(0.06, 0.14)), camber=float(rng.uniform(0.0, 0.04)),
                        position=float(rng.uniform(0.3, 0.5)), mach=float(rng.uniform(1.2, 2.0)),
                        aoa=float(rng.uniform(-2.0, 4.0))))
    return out


def build_one(spec: dict, preset: str = "LF2"):
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import config as CF
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import geometry as G
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import mesh as M
    poly, patches = G.reorder_points(G.naca4_points(spec["thickness"], spec["camber"], spec["position"], 100))
    case = CF.make_case(preset, {"freestream": {"Mach": spec["mach"], "AOA_degrees": spec["aoa"]}})
    mesh = M.build_case_mesh(case, poly, patches)
    params = M.params_of(case)
    return case, mesh, params, mesh.initialize(params, case.freestream)


SERVE_MIX = (("LF1", 0.50, 48_000), ("LF2", 0.35, 90_000), ("L7", 0.15, 180_000))  # preset, share, ~steps


def euler_lf1_workload(args) -> dict:
    """FP64 Euler, slip wall, HLLC, first order, SSP-RK3, CFL 0.15 on LF1 around a
    symmetric 12% NACA-4-style section (200 points), M 1.65, AoA 1.5 deg, farfield, to t* = 0.01,
    with SRD small cells and agglomerated.  The FP64 CPU oracle (one core) and the FP64 device
    session run the same case; steps must agree and the states within 1e-10 relative L2."""
    from oracle import pyoracle as O
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import config as CF
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import geometry as G
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import mesh as M
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import Session

    poly, patches = G.reorder_points(G.naca4_points(0.12, 0.0, 0.4, 100))
    rows = {}
    for label, agglomerate in (("srd", False), ("agglomerated", True)):
        case = CF.make_case("LF1", {"physics": {"model": "euler", "turbulence": "none"},
                                    "mesh": {"agglomerate_small": agglomerate}})
        mesh = M.build_case_mesh(case, poly, patches)
        params = M.params_of(case)
        u0 = mesh.initialize(params, case.freestream)
        t0 = time.perf_counter()
        ref, meta = O.run(mesh, params, u0, 0.01)
        cpu = time.perf_counter() - t0
        with Session(mesh, params, u0) as s:
            s.advance(0.01)  # warm-up solve (graph capture, first launches)
        with Session(mesh, params, u0) as s:
            t0 = time.perf_counter()
            s.advance(0.01)
            gpu = time.perf_counter() - t0
            got, gmeta = s.download(), dict(s.meta)
        nv = 4
        err = max(float(np.linalg.norm(got.reshape(nv, -1)[v] - ref.reshape(nv, -1)[v]) /
                        np.linalg.norm(ref.reshape(nv, -1)[v])) for v in range(nv))
        n = mesh.n_cells
        rows[label] = dict(cells=n, small_cells=int(mesh.info.small_cells), steps=meta["steps"],
                           gpu_steps=gmeta["steps"], cpu_seconds=round(cpu, 4), gpu_seconds=round(gpu, 4),
                           cpu_cell_steps_per_s=round(n * meta["steps"] / cpu), gpu_cell_steps_per_s=round(
                               n * gmeta["steps"] / gpu), state_relative_l2=err)
        log(f"Euler LF1 {label}: {n} cells, {meta['steps']} steps (device {gmeta['steps']}), CPU {cpu:.3f} s, "
            f"device {gpu:.4f} s, relative L2 {err:.2e}")
        if gmeta["steps"] != meta["steps"] or err > 1e-10:
            raise SystemExit(f"Euler LF1 {label}: device result differs from the oracle")
    cpu_rate = rows["agglomerated"]["cpu_cell_steps_per_s"]
    return {
        "metric": "FP64 Euler LF1 NACA 0012 to t*=0.01: cell-steps per second", "unit": "cell-steps/s",
        "value": rows["agglomerated"]["gpu_cell_steps_per_s"], "n_gpus": 1, "steps": rows["agglomerated"]["steps"],
        "warmup": 1, "ms_per_step": round(rows["agglomerated"]["gpu_seconds"] / rows["agglomerated"]["steps"] * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: symmetric 12% NACA-4-style section, uniform freestream start",
        "config": {"workload": "FP64 Euler LF1 NACA 0012, t*=0.01, SRD and agglomerated", "parallelism": "single GPU"},
        "cpu_baseline": {"value": cpu_rate, "unit": "cell-steps/s", "cores": 1, "kind": "port",
                         "sample": f"the full agglomerated run on the FP64 oracle ({rows['agglomerated']['steps']} steps)"},
        "detail": rows,
    }


def legacy_workload(args) -> dict:
    """The legacy tensor-product grid solver (include/ugrid_hip.h): --legacy-cases blast cases of
    --legacy-n x --legacy-n cells sharing one grid on [-1, 1]^2, transmissive sides, HLLC, first
    order, FP64, CFL 0.4; each case its own Riemann state pair (inside radius 0.3 / outside).
    --warmup untimed steps, then --steps steps timed (one device run call each way).
    roofline: the second RK stage kernel (reads the stage-1 candidate and the base state, writes
    the stage-2 candidate: 96 algorithmic bytes per cell) timed with HIP events per launch.
    cpu_baseline: the FP64 oracle of the legacy backend (one thread) on one case of 256 x 256."""
    from oracle import pygrid as PG
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import euler2d as E

    states = [(1.0, 0.0, 0.0, 1.0, 0.125, 0.0, 0.0, 0.1), (1.0, 0.0, 0.0, 1000.0, 1.0, 0.0, 0.0, 0.01),
              (1.0, 0.75, 0.0, 1.0, 0.125, 0.0, 0.0, 0.1), (1.0, 0.0, 0.3, 10.0, 1.0, 0.0, -0.2, 0.1)]

    def blast_batch(n, ncase, body=False):
        par = E.params2d_default()
        par.limiter = E.LIM_FIRST_ORDER
        s = E.Solver2D(n, n, -1.0, 1.0, -1.0, 1.0, ncase, par)
        X = s.grid.xc[2:2 + n][None, :]
        Y = s.grid.yc[2:2 + n][:, None]
        inside = np.broadcast_to(X * X + Y * Y < 0.09, (n, n))
        for c in range(ncase):
            st = states[c % len(states)]
            s.set_primitive_field(c, *[np.where(inside, st[k], st[4 + k]) for k in range(4)])
        if body:  # a solid disc of radius 0.2 beside the blast, as cut cells with SRD
            a = 0.0137 + 2.0 * np.pi * np.arange(256) / 256
            s.set_cut(E.cut2d_build_polygon(s, 0.55 + 0.2 * np.cos(a), 0.1 + 0.2 * np.sin(a)))
        return s

    n, ncase = args.legacy_n, args.legacy_cases
    cells = n * n * ncase
    s = blast_batch(n, ncase)
    log(f"legacy grid: {ncase} cases of {n}x{n} ({cells / 1e6:.1f}M cells), FP64 HLLC")
    with s.device_session() as d:
        info = d.launch_info()
        status, msg = d.run(1e9, args.warmup)
        assert status == 1, msg
        from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import _native as N
        t0 = time.perf_counter()
        status, msg = d.run(1e9, args.steps)
        wall = time.perf_counter() - t0
        assert status == 1, msg
        us = np.zeros(3)
        err = N.err_buffer()
        assert d.lib.ugrid_profile(d.handle, 20, N.dptr(us), err, N.ERR_LEN) == 0, err.value
        d.pull()
    assert np.all(np.isfinite(s.U))
    copy_rates = np.zeros(4)
    err = N.err_buffer()
    if N.load_library().amr_hip_copy_variants(1 << 30, 20, N.dptr(copy_rates), 4, err, N.ERR_LEN):
        raise SystemExit(err.value.decode())
    ms_per_step = wall / args.steps * 1e3
    value = cells * args.steps / wall
    stage_bytes = [64 * cells, 96 * cells, 96 * cells]
    rates = [stage_bytes[k] / (us[k] * 1e-6) / 1e9 for k in range(3)]
    log(f"legacy grid: {ms_per_step:.3f} ms/step, {value / 1e9:.2f} G cell-steps/s; stage us "
        f"{us[0]:.1f} / {us[1]:.1f} / {us[2]:.1f} -> {rates[0]:.0f} / {rates[1]:.0f} / {rates[2]:.0f} GB/s; "
        f"launch {info}")
    # The same batch with a solid body (cut cells, SRD after every stage): steps, rejected trials
    # and stage times including the redistribution kernels.
    cut = None
    if args.legacy_cut:
        sc = blast_batch(n, ncase, body=True)
        with sc.device_session() as d:
            status, msg = d.run(1e9, args.warmup)
            assert status == 1, msg
            r0 = d.meta.nreject
            t0 = time.perf_counter()
            status, msg = d.run(1e9, args.steps)
            wall_c = time.perf_counter() - t0
            assert status == 1, msg
            rejected = d.meta.nreject - r0
            d.pull()
        assert np.all(np.isfinite(sc.U))
        # Every step's first candidates fail the admissibility test in the smallest cut cells
        # (it runs before the redistribution, as in the reference's backend) until the step has
        # been halved enough: rejected trials run the first stage only.
        cut = {"cell_steps_per_second": round(cells * args.steps / wall_c), "ms_per_step": round(wall_c / args.steps * 1e3, 4),
               "rejected_trials": int(rejected), "ms_per_trial": round(wall_c / (args.steps + rejected) * 1e3, 4),
               "small_cells": int(sc.cut.nsmall), "cut_cells": int(sc.cut.ncut), "solid_cells": int(sc.cut.nsolid),
               "body": "disc of radius 0.2 at (0.55, 0.1), 256 vertices"}
        log(f"legacy grid with a body: {cut['ms_per_step']:.3f} ms/step ({rejected} rejected trials in "
            f"{args.steps} steps, {cut['ms_per_trial']:.3f} ms per trial), {cut['cell_steps_per_second'] / 1e9:.2f} "
            f"G cell-steps/s; {sc.cut.ncut} cut, {sc.cut.nsmall} small cells")
    # Backward Euler (pseudo-time defect relaxation) on the same batch at 512^2: physical steps,
    # pseudo updates (each one residual pass and one trial over every case) and their rate.
    be = None
    if args.legacy_be:
        nb = max(64, n // 2)
        sb = blast_batch(nb, ncase)
        sb.par.time_integrator = E.TIME_BACKWARD_EULER
        sb.par.cfl = 2.0
        with sb.device_session() as d:
            status, msg = d.run(1e9, 1)  # warm-up: one physical step
            assert status in (1, 4), msg
            st0 = d.implicit_stats()
            t0 = time.perf_counter()
            status, msg = d.run(1e9, args.legacy_be_steps)
            wall_b = time.perf_counter() - t0
            assert status in (1, 4), msg  # 4: a step not converged within the iteration limit
            st1 = d.implicit_stats()
        updates = st1["updates"] - st0["updates"]
        be = {"cells": nb * nb * ncase, "physical_steps": args.legacy_be_steps, "status": status,
              "pseudo_updates": int(updates),
              "retries": int(st1["retries"] - st0["retries"]), "seconds": round(wall_b, 3),
              "us_per_update": round(wall_b / max(updates, 1) * 1e6, 2),
              "cell_updates_per_second": round(nb * nb * ncase * updates / wall_b)}
        log(f"legacy grid backward Euler: {ncase} cases of {nb}x{nb}, {args.legacy_be_steps} physical steps, "
            f"{updates} pseudo updates in {wall_b:.2f} s ({be['us_per_update']:.1f} us each, "
            f"{be['cell_updates_per_second'] / 1e9:.2f} G cell-updates/s)")
    # CPU baseline: the oracle on one case of 256 x 256.
    cpu_n, cpu_steps = 256, 0
    c = blast_batch(cpu_n, 1)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < args.cpu_seconds:
        r = PG.run(c, 1e9, 5)
        c.U[:], c.t[:], c.dt[:] = r["u"], r["t"], r["dt"]
        cpu_steps += 5
    cpu_wall = time.perf_counter() - t0
    cpu_rate = cpu_n * cpu_n * cpu_steps / cpu_wall
    log(f"legacy grid cpu oracle (1 thread): {cpu_rate / 1e6:.2f}M cell-steps/s over {cpu_steps} steps")
    return {
        "metric": "legacy tensor-product grid solver: cell-steps per second (FP64, first order, SSP-RK3)",
        "value": round(value), "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic: circular blasts, four Riemann state pairs",
        "config": {"workload": f"legacy grid: {ncase} blast cases of {n}x{n}, HLLC, transmissive, CFL 0.4",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(rates[1], 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(rates[1] / 8000.0, 4), "traffic": None,
                     "kernel": "k_stage<2> (second RK stage: 96 algorithmic bytes per cell)",
                     "stage_us": [round(x, 2) for x in us], "stage_GBps": [round(x, 1) for x in rates],
                     "stream_copy_variants_gbps": [round(x, 1) for x in copy_rates]},
        "cpu_baseline": {"value": round(cpu_rate), "unit": "cell-steps/s", "cores": 1, "kind": "port",
                         "sample": f"{cpu_steps} steps of one {cpu_n}x{cpu_n} blast case on the FP64 oracle"},
        "launch": info,
        "detail": {"with_body": cut, "backward_euler": be},
    }


def serve_workload(args, n_requests: int | None = None) -> dict:
    """Continuous batching: requests (seeded airfoils on LF1 / LF2 / level-7 meshes, 50/35/15 %)
    arrive as a Poisson stream and are admitted into an arena of slot classes sized by the HBM
    budget; each runs to the steady criteria.  Reports p50/p99 time-to-steady and solves/hour."""
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import batch as B
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import config as CF
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import geometry as G
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import serve as S

    n_req = n_requests or args.serve
    rng = np.random.default_rng(args.seed)
    presets = rng.choice([p for p, _, _ in SERVE_MIX], size=n_req, p=[w for _, w, _ in SERVE_MIX])
    specs = draw_airfoils(n_req, args.seed + 1)
    requests = []
    for i, (spec, preset) in enumerate(zip(specs, presets)):
        case = CF.make_case(str(preset), {"freestream": {"Mach": spec["mach"], "AOA_degrees": spec["aoa"]}})
        poly = G.reorder_points(G.naca4_points(spec["thickness"], spec["camber"], spec["position"], 100))
        requests.append(S.Request(i, poly, case))
    # Slot classes: capacity of each preset from sample meshes (+5 %), counts proportional to the
    # expected residence (share x steps), within the HBM budget.
    caps, feats = {}, []
    for preset, _, _ in SERVE_MIX:
        sample = [build_one(s, preset)[1] for s in draw_airfoils(4, args.seed + 7)]
        cs, fs = zip(*[B.capacity(m, args.tile or 256) for m in sample])
        caps[preset] = B.envelope(cs, margin=0.05, tile=args.tile or 256)
        feats += fs
    weight = {p: w * st for p, w, st in SERVE_MIX}
    total = sum(weight.values())
    counts = {p: max(1, int(round(args.serve_slots * weight[p] / total))) for p in weight}
    slot_caps = [caps[p] for p, _, _ in SERVE_MIX for _ in range(counts[p])]
    features = B.union(feats)
    budget = int(args.hbm_budget_gb * 1e9)
    while B.arena_bytes(slot_caps, features, 32, args.tile or 256) > budget and len(slot_caps) > 3:
        slot_caps = slot_caps[: int(len(slot_caps) * 0.9)]
    arena = B.arena_bytes(slot_caps, features, 32, args.tile or 256)
    log(f"serving arena: {len(slot_caps)} slots {counts}, {arena / 1e9:.2f} GB of a {args.hbm_budget_gb:.0f} GB budget")
    backend = S.DeviceBackend(slot_caps, features, precision=32, tile=args.tile or 256)
    try:
        # Capacity estimate: step time of the arena filled with sample meshes.
        kinds = [p for p, _, _ in SERVE_MIX for _ in range(counts[p])][: len(slot_caps)]
        with ThreadPoolExecutor(args.workers) as pool:
            fill = list(pool.map(build_one, draw_airfoils(len(kinds), args.seed + 9), kinds))
        for slot, (case, mesh, params, u0) in enumerate(fill[: len(slot_caps)]):
            backend.batch.admit(slot, mesh, params, u0)
        backend.batch.set_targets(1e30)
        backend.batch.run(10)
        t0 = time.perf_counter()
        backend.batch.run(50)
        step = (time.perf_counter() - t0) / 50
        for slot in range(len(slot_caps)):
            backend.batch.release(slot)
        mean_steps = sum(w * st for _, w, st in SERVE_MIX)
        capacity = len(slot_caps) / (mean_steps * step)
        rate = args.serve_load * capacity
        log(f"arena step {step * 1e3:.3f} ms -> capacity ~{capacity:.3f} requests/s; arrival rate {rate:.3f}/s")
        arrivals = S.poisson_arrivals(rate, n_req, seed=args.seed + 3)
        for req, t in zip(requests, arrivals):
            req.arrival = float(t)
        server = S.Server(backend, mode="steady", budget=3600.0, workers=args.workers, poll_seconds=0.2,
                          log=lambda m: log(m))
        report = server.serve(requests)
    finally:
        backend.close()
    outcomes = report.pop("outcomes")
    lat = {p: [o.latency for o, r in zip(outcomes, requests) if not o.error and str(presets[r.rid]) == p]
           for p, _, _ in SERVE_MIX}
    return {
        "metric": METRIC, "value": round(report["solves_per_hour"], 2), "unit": "solves/hour", "n_gpus": 1,
        "steps": None, "warmup": None, "ms_per_step": round(step * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic: {n_req} seeded airfoils (seed {args.seed}), Poisson arrivals",
        "config": {"workload": f"serving: {n_req} steady LF1/LF2/L7 SA-neg requests, continuous batching",
                   "slots": len(slot_caps), "slot_classes": counts, "arena_bytes": arena,
                   "hbm_budget_gb": args.hbm_budget_gb, "arrival_rate_per_s": rate, "parallelism": "single GPU"},
        "detail": {**{k: (round(v, 4) if isinstance(v, float) else v) for k, v in report.items()},
                   "p50_by_preset": {p: S.percentile(v, 50) for p, v in lat.items()},
                   "p99_by_preset": {p: S.percentile(v, 99) for p, v in lat.items()}},
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--fp64-check", type=int, default=-1,
                    help="meshes re-solved in FP64 for dCL/dCD/dCM (-1: every mesh of the batch)")
    ap.add_argument("--few-slots", type=int, default=-1,
                    help="running slots at or below which the batch's flux launches take two threads per cell "
                         "(-1: autotune, 0: never)")
    ap.add_argument("--like-samples", type=int, default=1,
                    help="batch airfoils (seeded choice) also solved as plain FP64 single sessions")
    ap.add_argument("--serve-leg", type=int, default=48,
                    help="requests of the bounded serving leg that gives p50/p99 (0: none)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the threaded CPU baseline (0: OMP_NUM_THREADS, else the host's cores up to 16)")
    ap.add_argument("--budget", type=float, default=400.0, help="wall budget of each steady solve (s)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--solve", action=argparse.BooleanOptionalAction, default=True,
                    help="run the steady solves (--no-solve: timed steps, roofline and CPU baseline only)")
    ap.add_argument("--graph", action=argparse.BooleanOptionalAction, default=True,
                    help="replay captured hipGraphs in the timed steps (--no-graph: direct launches)")
    ap.add_argument("--tile", type=int, default=0, help="launch tile (0: autotune over 64/128/256)")
    ap.add_argument("--euler-lf1", action="store_true",
                    help="instead of the batch workload: FP64 Euler on LF1 to t*=0.01 on the oracle and the device")
    ap.add_argument("--serve", type=int, default=0,
                    help="instead of the batch workload: serve this many Poisson-arriving requests (LF1/LF2/L7 mix)")
    ap.add_argument("--serve-slots", type=int, default=192, help="slots of the serving arena")
    ap.add_argument("--serve-load", type=float, default=0.8, help="arrival rate as a fraction of estimated capacity")
    ap.add_argument("--hbm-budget-gb", type=float, default=200.0, help="device memory the serving arena may use")
    ap.add_argument("--legacy", action="store_true",
                    help="bench the legacy tensor-product grid solver instead (its own JSON line)")
    ap.add_argument("--legacy-n", type=int, default=1024, help="legacy grid: cells per side")
    ap.add_argument("--legacy-cases", type=int, default=16, help="legacy grid: cases in the batch")
    ap.add_argument("--legacy-cut", action=argparse.BooleanOptionalAction, default=True,
                    help="legacy grid: also time the batch with a solid body (cut cells, SRD)")
    ap.add_argument("--legacy-be", action=argparse.BooleanOptionalAction, default=True,
                    help="legacy grid: also time backward Euler on the batch at half the resolution")
    ap.add_argument("--legacy-be-steps", type=int, default=3, help="legacy grid: backward-Euler physical steps timed")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if os.environ.get("WORLD_SIZE", "1") not in ("", "1"):
        print("bench.py runs as a single process on one GPU", file=sys.stderr)
        sys.exit(2)

    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import _native as N
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import autotune as AT
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import batch as B
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import roofline as RL
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import Session, device_count

    if device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    lib = N.load_library()
    if args.euler_lf1:
        print(json.dumps(euler_lf1_workload(args)), flush=True)
        return
    if args.serve:
        print(json.dumps(serve_workload(args)), flush=True)
        return
    if args.legacy:
        print(json.dumps(legacy_workload(args)), flush=True)
        return

    # 0. Workload: seeded airfoils, LF2 meshes built on host threads (the mesher releases the GIL).
    specs = draw_airfoils(args.batch, args.seed)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(args.workers) as pool:
        items = list(pool.map(build_one, specs))
    mesh_seconds = time.perf_counter() - t0
    meshes = [m for _, m, _, _ in items]
    params = [p for _, _, p, _ in items]
    states = [u for _, _, _, u in items]
    cases = [c for c, _, _, _ in items]
    cells = [m.n_cells for m in meshes]
    log(f"{args.batch} LF2 meshes in {mesh_seconds:.1f} s: cells {min(cells)}..{max(cells)}, total {sum(cells)}")

    # 1. Tile autotuning on the batch (FP32).
    if args.tile:
        tile, tile_times = args.tile, {}
    else:
        tile, tile_times = AT.pick_tile(meshes, params, states, precision=32, trials=30, warmup=10)
    log("tile us/step: " + ", ".join(f"{t}: {s * 1e6:.1f}" for t, s in tile_times.items()) + f" -> {tile}")
    # Few-slot launch shape (two flux threads per cell once at most few_slots slots run).
    if args.few_slots < 0:
        few_slots, few_times = AT.pick_few_slots(meshes, params, states, precision=32, tile=tile)
        log("few-slot trial us (1 / 2 flux threads per cell): " +
            ", ".join(f"{k} running: {a} / {b}" for k, (a, b) in sorted(few_times.items())) + f" -> {few_slots}")
    else:
        few_slots, few_times = args.few_slots, {}

    # 2. Timed steps of the batch: exactly --steps graph replays after --warmup.
    with B.Batch.of_meshes(meshes, params, states, precision=32, tile=tile, graph=args.graph) as b:
        b.set_targets(1e30)
        b.run(args.warmup)  # returns after a device synchronisation (control read-back)
        t0 = time.perf_counter()
        b.run(args.steps)
        step_seconds = (time.perf_counter() - t0) / args.steps
        stats = b.stats()
        # Per-kernel device time over a few directly launched trials (hipEvent pairs on the arena stream).
        prof = b.profile(20)
    log(f"batch step {step_seconds * 1e6:.1f} us ({step_seconds * 1e6 / args.batch:.2f} us per mesh-step), "
        f"arena {stats['allocated_bytes'] / 1e6:.0f} MB, {stats['graph_nodes_per_trial']} graph nodes")
    per_launch = {k: v[0] / v[1] * 1e-3 for k, v in prof.items()}
    # The stage pass: the last RK stage's gradients and its fused faces + gather + RK update (the
    # first stage's gradient pass also opens the trial, so it is not a plain stage pass).
    stage_seconds = per_launch["gradient_stage"] + per_launch["flux_rk2"]
    b_design = sum(RL.design_stage_bytes(m, 32) for m in meshes)
    b_contract = sum(RL.contract_stage_bytes(m, 32) for m in meshes)
    achieved = b_contract / stage_seconds / 1e9
    achieved_design = b_design / stage_seconds / 1e9
    dominant = max(prof.items(), key=lambda kv: kv[1][0])[0]
    log("kernel us/launch: " + ", ".join(f"{k}={v * 1e6:.1f}" for k, v in per_launch.items()))
    log(f"stage pass {stage_seconds * 1e6:.1f} us: contract bytes {b_contract / 1e6:.1f} MB -> {achieved:.0f} GB/s "
        f"({achieved / RL.HBM_PEAK_GBPS:.1%} of 8 TB/s); this build's bytes {b_design / 1e6:.1f} MB -> "
        f"{achieved_design:.0f} GB/s")
    rates = (C.c_double * 4)()
    err = N.err_buffer()
    if lib.amr_hip_copy_variants(1 << 30, 20, rates, 4, err, N.ERR_LEN):
        raise SystemExit(f"copy probe: {err.value.decode()}")
    copy_rates = dict(zip(("1x16B", "4x16B", "4x16B_nontemporal", "8x16B_nontemporal"), (round(r, 1) for r in rates)))
    copy_best = max(rates)
    log("STREAM copy GB/s: " + ", ".join(f"{k} {v:.0f}" for k, v in copy_rates.items()))

    value, solve = None, {}
    single = build_one(dict(thickness=0.10, camber=0.02, position=0.4, mach=1.65, aoa=1.5))
    case1, mesh1, params1, u1 = single
    r64 = None
    if args.solve:
        # 3. Batched FP32 steady solve of the whole batch (the headline value).
        with B.Batch.of_meshes(meshes, params, states, precision=32, tile=tile) as b:
            if tile * 2 <= 256 and few_slots > 0:
                b.set_launch_shape(1, few_slots, 2)  # two flux threads per cell once few slots run
            t0 = time.perf_counter()
            res32 = b.solve(cases, mode="steady", budget=args.budget)
            wall32 = time.perf_counter() - t0
            solve_stats = b.stats()
        conv32 = sum(1 for r in res32 if r["converged"])
        steps32 = [r["steps"] for r in res32]
        value = 3600.0 * conv32 / wall32
        log(f"batched FP32 steady: {conv32}/{args.batch} converged in {wall32:.1f} s, steps "
            f"min/p50/mean/p90/max {min(steps32)}/{np.median(steps32):.0f}/{np.mean(steps32):.0f}/"
            f"{np.percentile(steps32, 90):.0f}/{max(steps32)} -> {value:.0f} solves/hour; "
            f"{solve_stats['status_polls']} control polls, {solve_stats['advance_seconds']:.1f} s advancing; "
            f"ideal at the full-batch step time {sum(steps32) * step_seconds / args.batch:.1f} s")

        # 4. FP64 re-solve (FP64 batch == FP64 sessions bitwise), untimed: every mesh by default.
        k = args.batch if args.fp64_check < 0 else min(args.fp64_check, args.batch)
        deltas = None
        if k:
            with B.Batch.of_meshes(meshes[:k], params[:k], states[:k], precision=64, tile=tile) as b:
                t0 = time.perf_counter()
                res64 = b.solve(cases[:k], mode="steady", budget=args.budget)
                wall64_batch = time.perf_counter() - t0
            ok = [i for i in range(k) if res64[i]["converged"] and res32[i]["converged"]]
            d = np.abs(np.array([[res32[i][q] - res64[i][q] for q in ("cl", "cd", "cm")] for i in ok]))
            deltas = dict(meshes=len(ok), of=k, fp64_batch_wall_seconds=round(wall64_batch, 2),
                          fp64_rejected_total=int(sum(r["rejected"] for r in res64[:k])),
                          fp64_steps_minus_fp32_max=int(max(abs(res64[i]["steps"] - res32[i]["steps"]) for i in ok))
                          if ok else None)
            for j, q in enumerate(("dcl", "dcd", "dcm")):
                deltas[f"max_abs_{q}"] = float(d[:, j].max()) if ok else None
                deltas[f"p99_abs_{q}"] = float(np.percentile(d[:, j], 99)) if ok else None
                deltas[f"meshes_within_1e-4_{q}"] = int((d[:, j] <= 1e-4).sum()) if ok else 0
            log(f"FP32 vs FP64 over {len(ok)} of {k} meshes ({wall64_batch:.1f} s FP64 batch): max |dCL| "
                f"{deltas['max_abs_dcl']:.2e} |dCD| {deltas['max_abs_dcd']:.2e} |dCM| {deltas['max_abs_dcm']:.2e}; "
                f"p99 {deltas['p99_abs_dcl']:.2e} / {deltas['p99_abs_dcd']:.2e} / {deltas['p99_abs_dcm']:.2e}")

        # 5. Plain path: one FP64 session (graph replays, 32 per poll) on the single LF2 case.
        with Session(mesh1, params1, u1, precision=64) as s:
            t0 = time.perf_counter()
            r64 = s.solve(case1, mode="steady", budget=args.budget, history=0)
            wall64 = time.perf_counter() - t0
        fp64_value = 3600.0 / wall64
        log(f"plain FP64 single solve: converged={r64['converged']} steps={r64['steps']} wall {wall64:.2f} s "
            f"-> {fp64_value:.0f} solves/hour; CL {r64['cl']:.9f} CD {r64['cd']:.9f} CM {r64['cm']:.9f}")
        # Like for like: the batch's own airfoils as plain FP64 single solves.  A seeded sample is
        # solved and timed; every mesh's FP64 step count is known from the FP64 batch (bitwise the
        # FP64 session's), so the whole batch's plain-path time is its steps times the measured
        # per-step time of the single sessions (the session's own tile may round the face sums in
        # another order than the batch's, so its step count can differ by a few).
        like = None
        singles = [dict(mesh="single-case airfoil", steps=int(r64["steps"]), wall_seconds=round(wall64, 3),
                        us_per_step=round(wall64 / r64["steps"] * 1e6, 2))]
        if deltas is not None and args.like_samples > 0:
            pick = np.random.default_rng(args.seed + 1).choice(k, size=min(args.like_samples, k), replace=False)
            for i in sorted(int(x) for x in pick):
                with Session(meshes[i], params[i], states[i], precision=64) as s:
                    t0 = time.perf_counter()
                    ri = s.solve(cases[i], mode="steady", budget=args.budget, history=0)
                    wi = time.perf_counter() - t0
                singles.append(dict(mesh=f"batch airfoil {i}", steps=int(ri["steps"]),
                                    fp64_batch_steps=int(res64[i]["steps"]), wall_seconds=round(wi, 3),
                                    us_per_step=round(wi / ri["steps"] * 1e6, 2)))
            us_step = float(np.mean([s_["us_per_step"] for s_ in singles]))
            steps64 = [res64[i]["steps"] for i in range(k) if res64[i]["converged"]]
            per_solve = float(np.mean(steps64)) * us_step * 1e-6
            like = dict(fp64_single_solves_per_hour=round(3600.0 / per_solve, 2),
                        speedup_vs_fp64_single=round(value / (3600.0 / per_solve), 2),
                        fp64_steps_mean=round(float(np.mean(steps64)), 1), meshes=len(steps64),
                        fp64_single_us_per_step=round(us_step, 2), timed_single_solves=singles)
            log(f"like for like: the batch's {len(steps64)} airfoils as plain FP64 single solves, mean "
                f"{np.mean(steps64):.0f} steps x {us_step:.2f} us/step (timed: " +
                ", ".join(f"{s_['mesh']} {s_['steps']} steps {s_['wall_seconds']:.2f} s" for s_ in singles) +
                f") -> {3600.0 / per_solve:.0f} solves/hour; batched FP32 / plain FP64 = "
                f"{value / (3600.0 / per_solve):.2f}x")
        with Session(mesh1, params1, u1, precision=32) as s:
            t0 = time.perf_counter()
            r32 = s.solve(case1, mode="steady", budget=args.budget, history=0)
            wall32_single = time.perf_counter() - t0
        log(f"FP32 single solve: steps={r32['steps']} wall {wall32_single:.2f} s, dCL {r32['cl'] - r64['cl']:.2e} "
            f"dCD {r32['cd'] - r64['cd']:.2e} dCM {r32['cm'] - r64['cm']:.2e}")
        solve = {
            "converged": conv32, "batch_wall_seconds": round(wall32, 2),
            "fp32_rejected_total": int(sum(r["rejected"] for r in res32)),
            "fp32_rejected_max": int(max(r["rejected"] for r in res32)),
            "steps_to_steady_min_mean_max": [int(min(steps32)), float(np.mean(steps32)), int(max(steps32))],
            "fp64_single_solves_per_hour": round(fp64_value, 2), "fp64_single_wall_seconds": round(wall64, 3),
            "fp64_single_steps": r64["steps"], "speedup_vs_fp64_single": round(value / fp64_value, 2),
            "fp32_vs_fp64_batch": deltas,
            "like_for_like": like,
            "single_case_fp64_cl_cd_cm": [r64["cl"], r64["cd"], r64["cm"]],
            "single_case_fp32_minus_fp64": [r32[q] - r64[q] for q in ("cl", "cd", "cm")],
            "single_case_fp32_wall_seconds": round(wall32_single, 3),
        }

    # 6. Serving leg (bounded): p50 / p99 time-to-steady under continuous batching.
    latency = None
    if args.solve and args.serve_leg > 0:
        args.tile = args.tile or tile
        leg = serve_workload(args, args.serve_leg)
        latency = {"p50_seconds": leg["detail"]["p50_latency_seconds"],
                   "p99_seconds": leg["detail"]["p99_latency_seconds"],
                   "solves_per_hour": leg["value"], "requests": args.serve_leg,
                   "converged": leg["detail"]["converged"], "failed": leg["detail"]["failed"],
                   "makespan_seconds": leg["detail"]["makespan_seconds"],
                   "p50_by_preset": leg["detail"]["p50_by_preset"], "p99_by_preset": leg["detail"]["p99_by_preset"],
                   "slots": leg["config"]["slots"], "arrival_rate_per_s": round(leg["config"]["arrival_rate_per_s"], 4),
                   "workload": leg["config"]["workload"]}

    # 7. CPU baseline: the FP64 oracle on the host, bounded samples of the single case.
    from oracle import pyoracle as O
    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or min(16, os.cpu_count() or 1)

    def cpu_sample(threaded: bool, seconds: float):
        os.environ["OMP_NUM_THREADS"] = str(threads)
        steps_cpu, t_cpu, u, tnow = 0, 0.0, u1, 0.0
        while True:  # at least one chunk of 20 steps, then until `seconds` of oracle time
            t0 = time.perf_counter()
            u, meta = O.run(mesh1, params1, u, 1e30, maxsteps=20, time=tnow, threaded=threaded)
            t_cpu += time.perf_counter() - t0
            steps_cpu += meta["steps"]
            tnow = meta["time"]
            if meta["steps"] < 20 or t_cpu >= seconds:
                return steps_cpu, t_cpu

    model = ""
    try:
        model = next(line.split(":", 1)[1].strip() for line in open("/proc/cpuinfo") if line.startswith("model name"))
    except (OSError, StopIteration):
        pass
    samples = {}
    for label, threaded, n_threads, seconds in (("single", False, 1, args.cpu_seconds),
                                                 ("threaded", True, threads, args.cpu_seconds * 0.75)):
        steps_cpu, t_cpu = cpu_sample(threaded, seconds)
        cpu_step = t_cpu / steps_cpu
        samples[label] = dict(threads=n_threads, steps=steps_cpu, seconds=round(t_cpu, 2),
                              ms_per_step=round(cpu_step * 1e3, 3))
        if r64 is not None:
            samples[label]["solves_per_hour"] = round(3600.0 / (r64["steps"] * cpu_step), 3)
        log(f"cpu oracle ({n_threads} thread{'s' if n_threads > 1 else ''}): {cpu_step * 1e3:.2f} ms/step over "
            f"{steps_cpu} steps ({t_cpu:.1f} s)")
    best = min(samples, key=lambda k: samples[k]["ms_per_step"])
    b = samples[best]
    scale = f", scaled by its {r64['steps']} FP64 steps to steady" if r64 is not None else ""
    cpu_baseline = {
        "value": b["solves_per_hour"] if r64 is not None else b["ms_per_step"],
        "unit": "solves/hour" if r64 is not None else "ms/step", "cores": b["threads"], "kind": "port",
        "sample": f"{b['steps']} FP64 oracle steps of the single LF2 case from the freestream start on "
                  f"{b['threads']} thread(s) ({b['seconds']} s, {b['ms_per_step']} ms/step){scale}",
        "host_cpu": model, "host_cores": os.cpu_count(), "single_thread": samples["single"],
        "threaded": samples["threaded"]}

    out = {
        "metric": METRIC,
        "value": round(value, 2) if value is not None else None,
        "unit": "solves/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_seconds * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: {args.batch} seeded NACA-4-style airfoils (seed {args.seed}), uniform freestream start",
        "config": {
            "workload": f"batched LF2 SA-neg steady solves, {args.batch} meshes in one arena, FP32 fluxes/work "
                        f"arrays with FP64 state, control and reductions",
            "batch": args.batch, "seed": args.seed, "mesh": "LF2 (6 levels, wall h/c 1/128)",
            "cells_total": int(sum(cells)), "cells_min": int(min(cells)), "cells_max": int(max(cells)),
            "tile": tile, "parallelism": "single GPU",
            "few_slot_flux_threads_per_cell": {"slots": few_slots, "threads": 2, "probe_us": few_times}
            if tile * 2 <= 256 and few_slots > 0 else None,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": RL.HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / RL.HBM_PEAK_GBPS, 4), "traffic": profiled_traffic(args.batch, tile, args.seed),
            "kernel": "stage pass of the batch: gradient_stage (k_grad<2>) + flux_rk2 (k_flux<2>, faces + gather + RK)",
            "bytes_per_launch": int(b_contract), "design_bytes_per_launch": int(b_design),
            "design_achieved": round(achieved_design, 1), "design_frac": round(achieved_design / RL.HBM_PEAK_GBPS, 4),
            "us_per_launch": round(stage_seconds * 1e6, 2), "stream_copy_gbps": round(copy_best, 1),
            "stream_copy_variants_gbps": copy_rates,
        },
        "latency": latency,
        "cpu_baseline": cpu_baseline,
        "detail": {
            **solve,
            "tile_us_per_step": {str(t): round(s * 1e6, 1) for t, s in tile_times.items()},
            "kernel_us_per_launch": {kname: round(v * 1e6, 2) for kname, v in per_launch.items()},
            "dominant_kernel": dominant, "arena_bytes": stats["allocated_bytes"],
            "mesh_seconds": round(mesh_seconds, 2),
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
