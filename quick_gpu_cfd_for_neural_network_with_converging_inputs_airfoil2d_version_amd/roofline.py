"""Algorithmic bytes of one SSP-RK3 stage of the cut-cell time step (the roofline contract).

One stage of one mesh = one residual evaluation at U^(s) plus one stage update.  The bytes a
perfectly fused stage must move to and from HBM, with wu the conserved-state width, ws the
width of the other per-cell work arrays, wg the geometry width, nv the conserved variables
(5 with SA-neg), g the gradient components kept between passes (4 laminar, 6 SA-neg, 0
Euler), int32 indices:

    B_stage = wu*3*nv*n                       read U^(s), read U^n, write U^(s+1)
            + ws*2*g*n                        gradients written and read once
            + nf*(12 + 6*wg)                  face left, right, axis|side; length, dx, dy, ex, ey, dist
            + 4*(ncsr + n + 1)                cell -> face CSR
            + n*(3*wg + 8)                    volume, dt area, wall distance; wall begin/count
            + nt*(4 + 2*wg) + 4*(n + 1)       least-squares terms (source + weights) and offsets
            + nw*(8 + 4*wg)                   wall owner, donor; nx, ny, length, donor distance
            + [base stage] 16*n               local dt write + read (FP64)
            + [SRD] wu*2*nv*n + 8*members + wg*n

The contract as first written has one state width (wu = ws = wg = 4 in FP32 mode, 8 in FP64
mode).  This build keeps the conserved state in FP64 in both modes (an FP32 state stalls above
the steady tolerance because the update dt*r falls below one ulp of u), so its FP32 mode moves
wu = 8, ws = wg = 4: `design_stage_bytes` counts those bytes, `contract_stage_bytes` the
single-width figure; the benchmark reports achieved bandwidth with the design bytes and shows
both.  A step is three stages.  The achieved bandwidth of a stage is B_stage / (time of the
kernels that make up the stage pass: gradient + flux).
"""
from __future__ import annotations

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, published


def stage_bytes(n: int, nf: int, ncsr: int, nt: int, nw: int, nv: int, laminar: bool, sa: bool, wu: int, ws: int,
                wg: int, base: bool = False, srd_members: int = 0) -> int:
    g = 6 if sa else (4 if laminar else 0)
    b = wu * 3 * nv * n + ws * 2 * g * n + nf * (12 + 6 * wg) + 4 * (ncsr + n + 1) + n * (3 * wg + 8)
    if laminar:
        b += nt * (4 + 2 * wg) + 4 * (n + 1)
    b += nw * (8 + 4 * wg)
    if base:
        b += 16 * n
    if srd_members:
        b += wu * 2 * nv * n + 8 * srd_members + wg * n
    return int(b)


def _counts(mesh):
    info, view = mesh.info, mesh.view
    small = int(info.small_cells) > 0
    return dict(n=int(info.n_cells), nf=int(info.n_faces), ncsr=int(info.n_cell_faces), nt=int(info.n_visc_terms),
                nw=int(info.n_walls), nv=5 if view.transport.sa_model else 4, laminar=view.transport.model == 1,
                sa=view.transport.sa_model == 1, srd_members=int(info.n_group_members) if small else 0)


def design_stage_bytes(mesh, precision: int, base: bool = False) -> int:
    """B_stage of a built mesh with this build's widths (state FP64; work and geometry FP32 or FP64)."""
    w = 8 if precision == 64 else 4
    return stage_bytes(**_counts(mesh), wu=8, ws=w, wg=w, base=base)


def contract_stage_bytes(mesh, precision: int, base: bool = False) -> int:
    """B_stage with one width for state, work and geometry (4 in FP32 mode, 8 in FP64 mode)."""
    w = 8 if precision == 64 else 4
    return stage_bytes(**_counts(mesh), wu=w, ws=w, wg=w, base=base)


def mesh_stage_bytes(mesh, precision: int, base: bool = False) -> int:
    return design_stage_bytes(mesh, precision, base)
