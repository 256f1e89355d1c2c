#!/usr/bin/env python3
"""Benchmark of the MI355X cut-cell time step: LF2 SA-neg steady solves per hour.

Workload (synthetic): seeded NACA-4-style section (thickness 10%, camber 2% at 0.4c), LF2
mesh, SA-neg RANS with Sutherland viscosity, HLLC, CFL 0.15 (viscous/SA 0.1), freestream
M 1.65, AoA 1.5 deg, started from uniform freestream.  A "step" is one SSP-RK3 time step of
the batch (hipGraph of one trial).  The timed region is exactly --steps steps after --warmup
untimed steps, bracketed by device synchronisation.

value = steady solves per hour = 3600 * batch / (steps_to_steady * seconds_per_step), with
steps_to_steady measured by running the same case to the steady criteria in this run.  The
plain FP64 single solve is measured the same way (detail.fp64_*); the CPU baseline is the FP64
oracle on the host on a bounded sample of steps of the same mesh.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "LF2 SA-neg steady solves/hour batched FP32 vs plain FP64; p50/p99 latency; HBM GB/s"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_case():
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import config as CF
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import geometry as G
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import mesh as M
    poly, patches = G.reorder_points(G.naca4_points(0.10, 0.02, 0.4, 100))
    case = CF.make_case("LF2")
    mesh = M.build_case_mesh(case, poly, patches)
    params = M.params_of(case)
    return case, mesh, params, mesh.initialize(params, case.freestream)


def timed_steps(session, steps: int, warmup: int) -> float:
    """Seconds per step over exactly `steps` steps (after `warmup`), device-synchronised."""
    if warmup:
        session.run(warmup)  # returns after a device sync (control read-back)
    t0 = time.perf_counter()
    session.run(steps)
    return (time.perf_counter() - t0) / steps


def steady_solve(mesh, params, u0, case, precision):
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import Session
    with Session(mesh, params, u0, precision=precision) as s:
        t0 = time.perf_counter()
        out = s.solve(case, mode="steady", budget=300.0, history=0)
        out["wall_seconds"] = time.perf_counter() - t0
    return out


def cpu_baseline(mesh, params, u0, seconds: float = 12.0):
    """FP64 oracle, one host thread, on a bounded sample of steps of the same mesh."""
    from oracle import pyoracle as O
    steps, t_total, u = 0, 0.0, u0
    chunk = 20
    time_now = 0.0
    while t_total < seconds:
        t0 = time.perf_counter()
        u_next, meta = O.run(mesh, params, u, 1e30, maxsteps=chunk, time=time_now)
        t_total += time.perf_counter() - t0
        steps += meta["steps"]
        u, time_now = u_next, meta["time"]
        if meta["steps"] < chunk:
            break
    return t_total / steps, steps, t_total


def kernel_profile(mesh, params, u0, precision, trials=40):
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import Session
    with Session(mesh, params, u0, precision=precision) as s:
        s.run(20)
        prof = s.profile(trials)
    return prof


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if os.environ.get("WORLD_SIZE", "1") not in ("", "1"):
        print("bench.py runs as a single process on one GPU", file=sys.stderr)
        sys.exit(2)

    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import _native as N
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import roofline as RL
    from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import Session, device_count
    import ctypes as C

    if device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    case, mesh, params, u0 = build_case()
    log(f"LF2 mesh: {mesh.n_cells} cells, {mesh.info.n_faces} faces, {mesh.info.n_walls} wall fragments")

    # 1. Timed steps, FP32 (the measured batch) and FP64 (the plain single-solve path).
    per_step = {}
    for prec in (args.precision, 64) if args.precision != 64 else (64,):
        with Session(mesh, params, u0, precision=prec) as s:
            per_step[prec] = timed_steps(s, args.steps, args.warmup)
        log(f"fp{prec}: {per_step[prec] * 1e6:.1f} us/step over {args.steps} steps")

    # 2. Steps to steady (full runs of the same case on the device, steady criteria of the reference).
    solves = {}
    for prec in sorted(per_step, reverse=True):
        solves[prec] = steady_solve(mesh, params, u0, case, prec)
        r = solves[prec]
        log(f"fp{prec} steady: converged={r['converged']} steps={r['steps']} t*={r['time']:.3g} "
            f"wall={r['wall_seconds']:.2f}s CL={r['cl']:.9f} CD={r['cd']:.9f} CM={r['cm']:.9f} "
            f"raw residual {max(r['residual']):.3e} ({r['stop_reason']})")

    prec = args.precision
    steps_to_steady = solves[prec]["steps"]
    value = 3600.0 * 1 / (steps_to_steady * per_step[prec])
    fp64_value = 3600.0 / (solves[64]["steps"] * per_step[64])

    # 3. Roofline of the stage pass (gradient + face + collect of one RK stage), live HIP events.
    prof = kernel_profile(mesh, params, u0, prec)
    stage_kernels = ["gradient_stage", "face_stage", "collect_rk1"]
    launches = prof["face_stage"][1] / 2
    stage_ms = (prof["gradient_stage"][0] / 2 + prof["face_stage"][0] / 2 + prof["collect_rk1"][0]) / launches
    b_stage = RL.mesh_stage_bytes(mesh, prec)
    achieved = b_stage / (stage_ms * 1e-3) / 1e9
    dominant = max(prof.items(), key=lambda kv: kv[1][0])
    log("kernel ms/launch: " + ", ".join(f"{k}={v[0] / v[1] * 1e3:.1f}us" for k, v in prof.items()))
    gbps, ms_copy = C.c_double(), C.c_double()
    err = N.err_buffer()
    N.load_library().amr_hip_copy_bandwidth(1 << 30, 20, C.byref(gbps), C.byref(ms_copy), err, N.ERR_LEN)
    log(f"stage pass {stage_ms * 1e3:.1f} us, {b_stage / 1e6:.2f} MB algorithmic -> {achieved:.1f} GB/s; "
        f"STREAM copy {gbps.value:.0f} GB/s")

    # 4. CPU baseline: FP64 oracle, one thread, bounded sample.
    cpu_s, cpu_steps, cpu_total = cpu_baseline(mesh, params, u0, args.cpu_seconds)
    cpu_value = 3600.0 / (solves[64]["steps"] * cpu_s)
    log(f"cpu oracle: {cpu_s * 1e3:.2f} ms/step over {cpu_steps} steps ({cpu_total:.1f} s)")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(per_step[prec] * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == 32 else "f64",
        "data": "synthetic (seeded NACA-4-style airfoil, uniform freestream start)",
        "config": {
            "workload": "LF2 SA-neg single steady solve, FP32 state (FP64 control/reductions)" if prec == 32
            else "LF2 SA-neg single steady solve, FP64",
            "mesh": "LF2 (levels 6, wall h/c 1/128)", "cells": mesh.n_cells, "batch": 1,
            "airfoil": "NACA-4-style t=0.10 m=0.02 p=0.4, 201 points", "mach": case.mach, "aoa_deg": case.aoa,
            "steps_to_steady": steps_to_steady, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": RL.HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / RL.HBM_PEAK_GBPS, 5), "traffic": None,
            "kernel": "stage pass (gradient_stage + face_stage + collect_rk1)",
            "bytes_per_launch": b_stage, "us_per_launch": round(stage_ms * 1e3, 3),
            "stream_copy_gbps": round(gbps.value, 1),
        },
        "cpu_baseline": {
            "value": round(cpu_value, 4), "unit": "solves/hour", "cores": 1, "kind": "port",
            "sample": f"{cpu_steps} FP64 oracle steps of the same LF2 mesh from the freestream start "
                      f"({cpu_total:.1f} s, {cpu_s * 1e3:.2f} ms/step), scaled by the FP64 steps to steady",
        },
        "detail": {
            "fp64_us_per_step": round(per_step[64] * 1e6, 2), "fp64_solves_per_hour": round(fp64_value, 3),
            "fp64_steps_to_steady": solves[64]["steps"],
            "fp64_coefficients": [solves[64]["cl"], solves[64]["cd"], solves[64]["cm"]],
            "coefficients": [solves[prec]["cl"], solves[prec]["cd"], solves[prec]["cm"]],
            "delta_cl_cd_cm_vs_fp64": [solves[prec][k] - solves[64][k] for k in ("cl", "cd", "cm")],
            "converged": [bool(solves[p]["converged"]) for p in sorted(solves)],
            "steady_wall_seconds": {str(p): round(solves[p]["wall_seconds"], 3) for p in solves},
            "dominant_kernel": dominant[0],
            "kernel_us_per_launch": {k: round(v[0] / v[1] * 1e3, 3) for k, v in prof.items()},
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
