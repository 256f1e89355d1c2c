"""Algorithmic bytes of one SSP-RK3 stage of the cut-cell time step (the roofline contract).

One stage of one mesh = one residual evaluation at U^(s) plus one stage update.  The bytes a
perfectly fused stage must move to and from HBM, with ws the state width (8 FP64, 4 FP32), wg
the geometry width, nv the conserved variables (5 with SA-neg), g the gradient components kept
between passes (4 laminar, 6 SA-neg, 0 Euler), int32 indices:

    B_stage = ws*3*nv*n                       read U^(s), read U^n, write U^(s+1)
            + ws*2*g*n                        gradients written and read once
            + nf*(12 + 6*wg)                  face left, right, axis|side; length, dx, dy, ex, ey, dist
            + 4*(ncsr + n + 1)                cell -> face CSR
            + n*(3*wg + 8)                    volume, dt area, wall distance; wall begin/count
            + nt*(4 + 2*wg) + 4*(n + 1)       least-squares terms (source + weights) and offsets
            + nw*(8 + 4*wg)                   wall owner, donor; nx, ny, length, donor distance
            + [base stage] 16*n               local dt write + read (FP64)
            + [SRD] ws*2*nv*n + 8*members + wg*n

A step is three stages.  The achieved bandwidth of a stage is B_stage / (time of the kernels
that make up the stage pass: gradient + face + collect).
"""
from __future__ import annotations

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, published


def stage_bytes(n: int, nf: int, ncsr: int, nt: int, nw: int, nv: int, laminar: bool, sa: bool, ws: int, wg: int,
                base: bool = False, srd_members: int = 0) -> int:
    g = 6 if sa else (4 if laminar else 0)
    b = ws * 3 * nv * n + ws * 2 * g * n + nf * (12 + 6 * wg) + 4 * (ncsr + n + 1) + n * (3 * wg + 8)
    if laminar:
        b += nt * (4 + 2 * wg) + 4 * (n + 1)
    b += nw * (8 + 4 * wg)
    if base:
        b += 16 * n
    if srd_members:
        b += ws * 2 * nv * n + 8 * srd_members + wg * n
    return int(b)


def mesh_stage_bytes(mesh, precision: int, base: bool = False) -> int:
    """B_stage of a built mesh (exact element counts from the mesher)."""
    info, view = mesh.info, mesh.view
    ws = 8 if precision == 64 else 4
    small = int(info.small_cells) > 0
    return stage_bytes(int(info.n_cells), int(info.n_faces), int(info.n_cell_faces), int(info.n_visc_terms),
                       int(info.n_walls), 5 if view.transport.sa_model else 4, view.transport.model == 1,
                       view.transport.sa_model == 1, ws, ws, base, int(info.n_group_members) if small else 0)
