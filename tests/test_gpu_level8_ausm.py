"""Level-8 AUSM+-up laminar case (constant viscosity, pressure-outlet xmax, M 0.75, AoA 2 deg,
seeded 14%-thick cambered airfoil, ~104k cells).

FP64 device path against the FP64 oracle over 30 steps (1e-10 relative L2 per variable, equal
step counts); FP32 against FP64 on the device to t* = 0.5: CL / CD / CM_z within 1e-4 and the
wall pressure coefficient within 1e-4 relative L2 (Cp = (p_owner - p_inf) / q_inf on every
wall fragment, q_inf = M^2 / 2 in reference units).
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import Session

from .common import case_mesh, naca_polygon, per_variable_l2, relative_l2

pytestmark = pytest.mark.gpu

OVER = {"freestream": {"Mach": 0.75, "AOA_degrees": 2.0}, "physics": {"turbulence": "none", "viscosity": "constant"},
        "solver": {"flux": "ausm"}, "boundaries": {"xmax": "pressure-outlet"}}


def level8_ausm_case():
    rng = np.random.default_rng(4)
    camber, position = rng.uniform(0.02, 0.04), rng.uniform(0.3, 0.5)
    case, mesh, params = case_mesh(naca_polygon(0.14, camber, position), "L8", OVER)
    return case, mesh, params, mesh.initialize(params, case.freestream)


def wall_cp(mesh, u, case):
    n = mesh.n_cells
    rho, mx, my, e = (u[i * n:(i + 1) * n] for i in range(4))
    p = (case.gamma - 1) * (e - 0.5 * (mx * mx + my * my) / rho)
    owner = mesh.arrays()["walls"]["owner_cell"].astype(np.int64)
    return (p[owner] - 1.0 / case.gamma) / case.q_inf


def test_level8_ausm_fp64_matches_oracle_over_30_steps():
    case, mesh, params, u0 = level8_ausm_case()
    print("L8 cells", mesh.n_cells, flush=True)
    ref, meta = O.run(mesh, params, u0, 1e30, maxsteps=30)
    assert meta["steps"] == 30
    with Session(mesh, params, u0) as s:
        s.run(30)
        got = s.download()
        assert s.meta["steps"] == 30 and s.meta["rejected"] == 0
    errs = per_variable_l2(got, ref, 4)
    print("relative L2 per variable", errs, flush=True)
    assert max(errs) <= 1e-10


def test_level8_ausm_fp32_against_fp64():
    case, mesh, params, u0 = level8_ausm_case()
    frame = dict(region=(-1.0, 3.0, -1.0, 1.0), moment=(0.25, 0.0), mach=case.mach, aoa=case.aoa)
    out = {}
    for prec in (64, 32):
        with Session(mesh, params, u0, precision=prec) as s:
            assert s.advance(0.5) == 0
            out[prec] = (s.download(), s.monitor(**frame), dict(s.meta))
    (u64, m64, meta64), (u32, m32, meta32) = out[64], out[32]
    d = {q: m32[q] - m64[q] for q in ("cl", "cd", "cm")}
    cp_err = relative_l2(wall_cp(mesh, u32, case), wall_cp(mesh, u64, case))
    print("steps", meta64["steps"], meta32["steps"], "dCL/dCD/dCM", d, "wall Cp rel L2", cp_err,
          "CL", m64["cl"], flush=True)
    assert all(abs(v) <= 1e-4 for v in d.values())
    assert cp_err <= 1e-4
