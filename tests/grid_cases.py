"""Problems of the tensor-product grid solver for the tests: the reference CLI's problem set
(src/euler2d_cli.cpp:382-431: isentropic vortex, Riemann problems along x / y with one case per
state pair, circular blast) plus seeded smooth random states."""
from __future__ import annotations

import math

import numpy as np

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import euler2d as E

SOD = (1.0, 0.0, 0.0, 1.0, 0.125, 0.0, 0.0, 0.1)
RIEMANN_SET = [
    SOD,
    (1.0, 0.75, 0.0, 1.0, 0.125, 0.0, 0.0, 0.1),          # modified Sod
    (1.0, -2.0, 0.0, 0.4, 1.0, 2.0, 0.0, 0.4),            # 123 problem (two rarefactions)
    (1.0, 0.0, 0.3, 1000.0, 1.0, 0.0, -0.2, 0.01),        # strong blast, shear across the contact
    (5.99924, 19.5975, 0.0, 460.894, 5.99242, -6.19633, 0.0, 46.0950),  # colliding shocks
    (1.4, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0),             # stationary contact
]


def params(flux=E.FLUX_HLLC, bc=(E.BC_TRANSMISSIVE,) * 4, cfl=0.4, dt_fixed=0.0, mco=1.0, sx=None, sy=None,
           gamma=1.4) -> E.Params2D:
    p = E.params2d_default()
    p.limiter = E.LIM_FIRST_ORDER
    p.flux = flux
    p.bc = list(bc)
    p.cfl = cfl
    p.dt_fixed = dt_fixed
    p.ausm_mco = mco
    p.gamma = gamma
    if sx is not None:
        p.sx = E.Stretch(*sx)
    if sy is not None:
        p.sy = E.Stretch(*sy)
    return p


def vortex(n=64, ny=None, par=None, beta=5.0, u0=1.0, v0=1.0, half=8.0, ncase=1) -> E.Solver2D:
    """Periodic Shu vortex (src/euler2d_cli.cpp:412-419)."""
    par = par or params(bc=(E.BC_PERIODIC,) * 4)
    par.bc = [E.BC_PERIODIC] * 4
    ny = ny or n
    s = E.Solver2D(n, ny, -half, half, -half, half, ncase, par)
    g = s.grid
    gm = par.gamma
    x = g.xc[2:2 + n][None, :]
    y = g.yc[2:2 + ny][:, None]
    for c in range(ncase):
        dx, dy = x - 0.3 * c, y + 0.2 * c
        e = np.exp(0.5 * (1.0 - (dx * dx + dy * dy)))
        u = u0 - beta / (2.0 * math.pi) * dy * e
        v = v0 + beta / (2.0 * math.pi) * dx * e
        T = 1.0 - (gm - 1.0) * beta * beta / (8.0 * gm * math.pi * math.pi) * e * e
        rho = T ** (1.0 / (gm - 1.0))
        s.set_primitive_field(c, rho, u, v, rho ** gm)
    return s


def riemann(nx=100, ny=8, axis="x", states=RIEMANN_SET[:3], par=None, x0=0.5) -> E.Solver2D:
    """One case per state pair; discontinuity at x0 along x (riemann_x) or y (riemann_y)."""
    par = par or params()
    s = E.Solver2D(nx, ny, 0.0, 1.0, 0.0, 1.0, len(states), par)
    g = s.grid
    X = np.broadcast_to(g.xc[2:2 + nx][None, :], (ny, nx))
    Y = np.broadcast_to(g.yc[2:2 + ny][:, None], (ny, nx))
    left = (X < x0) if axis == "x" else (Y < x0)
    for c, st in enumerate(states):
        L, R = np.array(st[:4]), np.array(st[4:])
        w = [np.where(left, L[k], R[k]) for k in range(4)]
        s.set_primitive_field(c, *w)
    return s


def blast(n=48, par=None, r0=0.3, states=(SOD,)) -> E.Solver2D:
    par = par or params(bc=(E.BC_REFLECT,) * 4)
    s = E.Solver2D(n, n, -1.0, 1.0, -1.0, 1.0, len(states), par)
    g = s.grid
    X = g.xc[2:2 + n][None, :]
    Y = g.yc[2:2 + n][:, None]
    inside = np.broadcast_to(X * X + Y * Y < r0 * r0, (n, n))
    for c, st in enumerate(states):
        w = [np.where(inside, st[k], st[4 + k]) for k in range(4)]
        s.set_primitive_field(c, *w)
    return s


def random_smooth(nx, ny, ncase, par, seed=0, amplitude=0.2) -> E.Solver2D:
    """Smooth positive states (a few Fourier modes) in every case; ghosts hold garbage."""
    rng = np.random.default_rng(seed)
    s = E.Solver2D(nx, ny, 0.0, 1.3, -0.4, 0.6, ncase, par)
    g = s.grid
    X = g.xc[2:2 + nx][None, :]
    Y = g.yc[2:2 + ny][:, None]
    for c in range(ncase):
        def mode():
            a, b, kx, ky = rng.uniform(-1, 1, 4)
            return amplitude * a * np.sin(2 * math.pi * (1 + abs(kx)) * X + 3 * b) * np.cos(2 * math.pi * (1 + abs(ky)) * Y)
        rho = 1.0 + mode()
        u = rng.uniform(-0.8, 0.8) + mode()
        v = rng.uniform(-0.8, 0.8) + mode()
        p = 1.0 + mode()
        s.set_primitive_field(c, rho, u, v, p)
        s.mco[c] = rng.uniform(0.05, 1.0)
    ghost = np.ones(s.U.shape, dtype=bool)
    q = ghost.reshape(ncase, 4, g.my, g.mx)
    q[:, :, 2:2 + ny, 2:2 + nx] = False
    s.U[ghost] = rng.uniform(-5, 5, int(ghost.sum()))
    return s


ALL_BCS = [
    (E.BC_PERIODIC, E.BC_PERIODIC, E.BC_PERIODIC, E.BC_PERIODIC),
    (E.BC_TRANSMISSIVE,) * 4,
    (E.BC_REFLECT,) * 4,
    (E.BC_REFLECT, E.BC_TRANSMISSIVE, E.BC_PERIODIC, E.BC_PERIODIC),
    (E.BC_TRANSMISSIVE, E.BC_REFLECT, E.BC_REFLECT, E.BC_TRANSMISSIVE),
    (E.BC_PERIODIC, E.BC_PERIODIC, E.BC_TRANSMISSIVE, E.BC_REFLECT),
]
STRETCHES = [(None, None), ((E.STRETCH_GEOMETRIC, 1.03), (E.STRETCH_TANH_ENDS, 1.5)),
             ((E.STRETCH_SINH_CENTER, 2.0), (E.STRETCH_GEOMETRIC, 0.97))]


# ------------------------------------------------------------------ cut-cell bodies --
def ellipse(cx, cy, rx, ry, n=64, phase=0.0137):
    """Vertices of an ellipse, counter-clockwise (a phase keeps vertices off grid lines)."""
    a = phase + 2.0 * math.pi * np.arange(n) / n
    return cx + rx * np.cos(a), cy + ry * np.sin(a)


def wedge():
    """A triangle with one vertex pointing upstream."""
    return np.array([-0.45, 0.357, 0.357]), np.array([0.013, -0.313, 0.341])


BODIES = {
    "circle": lambda: ellipse(0.1, 0.05, 0.33, 0.27),
    "ellipse": lambda: ellipse(-0.07, 0.0, 0.52, 0.11, 96),
    "wedge": wedge,
}


def cut_body(body="circle", nx=48, ny=40, par=None, states=((1.0, 0.6, 0.05, 1.0),), shock=None):
    """Flow past a solid body on [-1, 1] x [-0.8, 0.8]: one case per uniform state (rho, u, v, p),
    optionally with a normal shock upstream at x = shock (post-shock state left of it).  The
    body's cut geometry is built and attached; returns (solver, px, py)."""
    par = par or params(bc=(E.BC_TRANSMISSIVE, E.BC_TRANSMISSIVE, E.BC_REFLECT, E.BC_REFLECT))
    s = E.Solver2D(nx, ny, -1.0, 1.0, -0.8, 0.8, len(states), par)
    g = s.grid
    X = np.broadcast_to(g.xc[2:2 + nx][None, :], (ny, nx))
    for c, st in enumerate(states):
        w = [np.full((ny, nx), float(st[k])) for k in range(4)]
        if shock is not None:
            pre = w[0], w[1], w[3]
            left = X < shock
            w[0] = np.where(left, 2.0 * pre[0], pre[0])
            w[1] = np.where(left, 0.5 * pre[1] + 0.4, pre[1])
            w[3] = np.where(left, 2.6 * pre[2], pre[2])
        s.set_primitive_field(c, *w)
    px, py = BODIES[body]()
    s.set_cut(E.cut2d_build_polygon(s, px, py))
    return s, px, py
