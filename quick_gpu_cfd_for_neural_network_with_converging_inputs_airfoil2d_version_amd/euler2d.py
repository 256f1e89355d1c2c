"""Host side of the tensor-product grid Euler solver (the reference's legacy 2-D family,
src/euler2d.h:1-225) over the HIP backend of include/ugrid_hip.h.

Mirrors the reference interface: `params2d_default`, `grid2d_stretch_faces`, `Solver2D`
(`solver2d_init`'s grid and arrays, `set_primitive`, `get_primitive`, `set_mco`, `integral`,
`u2idx`), `cut2d_build_polygon` / `set_cut` (cut-cell geometry and SRD neighbourhoods of a
solid polygon, built natively by ugrid_cut_build), `solver2d_run_cuda` / `solver2d_run_backend`
/ `solver2d_cuda_available` with the reference's return codes (0 every case at tend, 1
maxsteps, 2 nonphysical, 3 unsupported or unavailable, 4 backward-Euler step not converged).
The device run covers what the reference's GPU backend covers: FP64, first order (limiter 3),
HLLC / AUSM+-up, SSP-RK3 or backward Euler by pseudo-time defect relaxation, regular or
cut-cell grids (state redistribution after every RK stage), per-case time, step and AUSM+-up
reference Mach number.  There is no CPU solver here: backend 0 is the reference's own CPU code
and is refused.
"""
from __future__ import annotations

import ctypes as C
import sys
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

NVAR2D = 4
FLUX_HLLC, FLUX_AUSMPUP = 0, 1
LIM_NONE, LIM_MINMOD, LIM_VENKAT, LIM_FIRST_ORDER = 0, 1, 2, 3
BC_PERIODIC, BC_TRANSMISSIVE, BC_REFLECT = 0, 1, 2
STRETCH_UNIFORM, STRETCH_GEOMETRIC, STRETCH_TANH_ENDS, STRETCH_SINH_CENTER = 0, 1, 2, 3
TIME_SSP_RK3, TIME_BACKWARD_EULER = 0, 1
BACKEND_CPU, BACKEND_CUDA = 0, 1


@dataclass
class Stretch:
    type: int = STRETCH_UNIFORM
    param: float = 0.0


@dataclass
class Params2D:
    """src/euler2d.h:38-52; defaults of params2d_default (src/euler2d.cpp:13-37)."""
    gamma: float = 1.4
    cfl: float = 0.4
    flux: int = FLUX_HLLC
    limiter: int = LIM_VENKAT
    venkat_k: float = 5.0
    ausm_mco: float = 1.0
    bc: list = field(default_factory=lambda: [BC_TRANSMISSIVE] * 4)
    dt_fixed: float = 0.0
    time_integrator: int = TIME_SSP_RK3
    pseudo_cfl: float = 0.4
    implicit_maxiter: int = 2000
    implicit_rtol: float = 1e-8
    implicit_atol: float = 1e-11
    sx: Stretch = field(default_factory=Stretch)
    sy: Stretch = field(default_factory=Stretch)

    def to_c(self) -> N.ugrid_params:
        return N.ugrid_params(self.gamma, self.cfl, self.flux, self.limiter, self.ausm_mco,
                              (C.c_int32 * 4)(*self.bc), self.dt_fixed, self.time_integrator, self.implicit_maxiter,
                              self.pseudo_cfl, self.implicit_rtol, self.implicit_atol)


def params2d_default() -> Params2D:
    return Params2D()


def _axis(n: int, a: float, b: float, s: Stretch, bc_lo: int, bc_hi: int):
    xf, xc, dx = np.zeros(n + 1), np.zeros(n + 4), np.zeros(n + 4)
    err = N.err_buffer()
    if N.load_library().ugrid_axis(n, a, b, s.type, s.param, bc_lo, bc_hi, N.dptr(xf), N.dptr(xc), N.dptr(dx), err,
                                   N.ERR_LEN):
        raise ValueError(err.value.decode())
    return xf, xc, dx


def grid2d_stretch_faces(n: int, a: float, b: float, s: Stretch) -> np.ndarray:
    """n + 1 face coordinates on [a, b], both ends exact (src/euler2d.cpp:38-62)."""
    return _axis(n, a, b, s, BC_TRANSMISSIVE, BC_TRANSMISSIVE)[0]


@dataclass
class Grid2D:
    nx: int
    ny: int
    xmin: float
    xmax: float
    ymin: float
    ymax: float
    xf: np.ndarray
    yf: np.ndarray
    xc: np.ndarray
    dx: np.ndarray
    yc: np.ndarray
    dy: np.ndarray
    ng: int = 2

    @property
    def mx(self) -> int:
        return self.nx + 2 * self.ng

    @property
    def my(self) -> int:
        return self.ny + 2 * self.ng


_CUT_CELL = ("vol", "cx", "cy", "flag", "wax", "way", "wcx", "wcy", "nbrcount", "smallpos")
_CUT_SMALL = ("sidx", "scnt", "svhat", "sxhat", "syhat")


class Cut2D:
    """Cut-cell geometry and SRD neighbourhoods of a solid polygon on a Solver2D's grid (the
    reference's Cut2D, src/euler2d.h:98-117): numpy copies of every array and the C view the
    device backend reads."""

    def __init__(self, s: "Solver2D", px, py, patches=None):
        g = s.grid
        px = np.ascontiguousarray(px, dtype=np.float64)
        py = np.ascontiguousarray(py, dtype=np.float64)
        pat = None if patches is None else np.ascontiguousarray(patches, dtype=np.int32)
        keep = [np.ascontiguousarray(a) for a in (g.xf, g.xc, g.dx, g.yf, g.yc, g.dy)]
        axes = N.ugrid_axes(g.nx, g.ny, *[N.dptr(a) for a in keep])
        lib = N.load_library()
        handle = C.c_void_p()
        err = N.err_buffer()
        self.status = lib.ugrid_cut_build(C.byref(axes), N.dptr(px), N.dptr(py), len(px),
                                          None if pat is None else pat.ctypes.data_as(N.c_i32_p), C.byref(handle),
                                          err, N.ERR_LEN)
        if self.status:
            raise ValueError(f"cut2d_build_polygon failed ({self.status}): {err.value.decode()}")
        try:
            v = N.ugrid_cut_view()
            lib.ugrid_cut_get(handle, C.byref(v))
            cells, nfx, nfy = g.mx * g.my, (g.nx + 1) * g.ny, g.nx * (g.ny + 1)
            self.nsmall, self.nbad_nbhd, self.nsolid, self.ncut = v.nsmall, v.nbad_nbhd, v.nsolid, v.ncut
            for name in _CUT_CELL:
                setattr(self, name, N.as_array(getattr(v, name), cells, np.float64))
            for name, count in (("axl", nfx), ("axy", nfx), ("ayl", nfy), ("ayx", nfy)):
                setattr(self, name, N.as_array(getattr(v, name), count, np.float64))
            for name in _CUT_SMALL:
                setattr(self, name, N.as_array(getattr(v, name), v.nsmall, np.float64))
            self.smem = N.as_array(v.smem, 25 * v.nsmall, np.float64)
            self.walls = N.as_array(v.walls, v.n_walls, N.WALL_DTYPE)
        finally:
            lib.ugrid_cut_destroy(handle)

    def view(self) -> N.ugrid_cut_view:
        """The C view of these arrays (valid while this object lives)."""
        v = N.ugrid_cut_view()
        v.nsmall, v.nbad_nbhd, v.nsolid, v.ncut = self.nsmall, self.nbad_nbhd, self.nsolid, self.ncut
        v.n_walls = 0
        for name in _CUT_CELL + ("axl", "axy", "ayl", "ayx") + _CUT_SMALL + ("smem",):
            setattr(v, name, N.dptr(getattr(self, name)))
        return v


def cut2d_build_polygon(s: "Solver2D", px, py, patches=None) -> Cut2D:
    """src/euler2d.h:167-170: raises ValueError with the reference's code on failure."""
    return Cut2D(s, px, py, patches)


class Solver2D:
    """solver2d_init (src/euler2d.cpp:107-151): the grid of two stretched axes with two ghost
    layers, U[((case * 4 + var) * my + j + 2) * mx + i + 2] zeroed, per-case t, dt and mco."""

    def __init__(self, nx: int, ny: int, xmin: float, xmax: float, ymin: float, ymax: float, ncase: int,
                 par: Params2D):
        if nx < 1 or ny < 1 or ncase < 1:
            raise ValueError("nx, ny and ncase must be positive")
        self.par = par
        self.ncase = ncase
        xf, xc, dx = _axis(nx, xmin, xmax, par.sx, par.bc[0], par.bc[1])
        yf, yc, dy = _axis(ny, ymin, ymax, par.sy, par.bc[2], par.bc[3])
        self.grid = Grid2D(nx, ny, xmin, xmax, ymin, ymax, xf, yf, xc, dx, yc, dy)
        g = self.grid
        self.U = np.zeros(ncase * NVAR2D * g.mx * g.my)
        self.t = np.zeros(ncase)
        self.dt = np.zeros(ncase)
        self.mco = np.full(ncase, par.ausm_mco, dtype=np.float64)
        self.nsteps = 0
        self.nreject = 0
        # ImplicitStats2D (src/euler2d.h:118-121): backward-Euler updates / retries (cumulative)
        # and the latest convergence check's largest first norm, norm and tolerance
        self.implicit = dict(updates=0, retries=0, initial_defect=0.0, final_defect=0.0, tolerance=0.0)
        self.last_error = ""
        self.cut = None  # Cut2D (solver2d_set_cut)

    def set_cut(self, cut: "Cut2D | None") -> None:
        """solver2d_set_cut (src/euler2d.h:164-165)."""
        self.cut = cut

    # -- field access (src/euler2d.h:137-140, src/euler2d.cpp:158-191) --
    def u2idx(self, c: int, v: int, i: int, j: int) -> int:
        g = self.grid
        return ((c * NVAR2D + v) * g.my + (j + g.ng)) * g.mx + (i + g.ng)

    def field(self, c: int) -> np.ndarray:
        """View of case c as (4, my, mx), ghosts included."""
        g = self.grid
        return self.U.reshape(self.ncase, NVAR2D, g.my, g.mx)[c]

    def set_primitive(self, c: int, i, j, rho, u, v, p) -> None:
        """Scalar or array (i, j index arrays / slices) form of solver2d_set_primitive."""
        g = self.grid
        q = self.field(c)
        gm = self.par.gamma
        rho, u, v, p = (np.asarray(a, dtype=np.float64) for a in (rho, u, v, p))
        jj, ii = np.asarray(j) + g.ng, np.asarray(i) + g.ng
        q[0, jj, ii] = rho
        q[1, jj, ii] = rho * u
        q[2, jj, ii] = rho * v
        q[3, jj, ii] = p / (gm - 1.0) + 0.5 * rho * u * u + 0.5 * rho * v * v

    def set_primitive_field(self, c: int, rho, u, v, p) -> None:
        """Every interior cell of case c from (ny, nx) arrays (or scalars)."""
        g = self.grid
        shape = (g.ny, g.nx)
        rho, u, v, p = (np.broadcast_to(np.asarray(a, dtype=np.float64), shape) for a in (rho, u, v, p))
        q = self.field(c)
        gm = self.par.gamma
        inner = (slice(g.ng, g.ng + g.ny), slice(g.ng, g.ng + g.nx))
        q[0][inner] = rho
        q[1][inner] = rho * u
        q[2][inner] = rho * v
        q[3][inner] = p / (gm - 1.0) + 0.5 * rho * u * u + 0.5 * rho * v * v

    def get_primitive(self, c: int, i, j):
        g = self.grid
        q = self.field(c)
        jj, ii = np.asarray(j) + g.ng, np.asarray(i) + g.ng
        r, mu, mv, E = (q[k, jj, ii] for k in range(4))
        p = (self.par.gamma - 1.0) * (E - 0.5 * mu * mu / r - 0.5 * mv * mv / r)
        return r, mu / r, mv / r, p

    def set_mco(self, c: int, mco: float) -> None:
        self.mco[c] = mco

    def integral(self, c: int, var: int) -> float:
        """sum U dx dy over the interior, in the reference's loop order (src/euler2d.cpp:193-203)."""
        g = self.grid
        q = self.field(c)[var]
        total = 0.0
        for j in range(g.ny):
            row = q[j + g.ng, g.ng:g.ng + g.nx]
            for i in range(g.nx):
                total += row[i] * (g.dx[i + g.ng] * g.dy[j + g.ng])
        return total

    # -- device backend --
    def _setup(self) -> tuple:
        g = self.grid
        dx = np.ascontiguousarray(g.dx)
        dy = np.ascontiguousarray(g.dy)
        mco = np.ascontiguousarray(self.mco, dtype=np.float64)
        cut = self.cut.view() if self.cut is not None else None
        setup = N.ugrid_setup(g.nx, g.ny, self.ncase, 0, N.dptr(dx), N.dptr(dy), self.par.to_c(), N.dptr(mco),
                              C.pointer(cut) if cut is not None else None)
        return setup, (dx, dy, mco, cut, self.cut)

    def device_session(self) -> "DeviceSession":
        return DeviceSession(self)


class DeviceSession:
    """A batch resident on the device (ugrid_create ... ugrid_destroy).  run() advances it;
    pull() copies state, times and last steps back into the Solver2D."""

    def __init__(self, s: Solver2D):
        self.lib = N.load_library()
        self.solver = s
        setup, self._keep = s._setup()
        self.handle = C.c_void_p()
        err = N.err_buffer()
        u = np.ascontiguousarray(s.U)
        t = np.ascontiguousarray(s.t, dtype=np.float64)
        dt = np.ascontiguousarray(s.dt, dtype=np.float64)
        status = self.lib.ugrid_create(C.byref(setup), N.dptr(u), u.size, N.dptr(t), N.dptr(dt),
                                       C.byref(self.handle), err, N.ERR_LEN)
        if status:
            raise RuntimeError(err.value.decode())
        self.meta = N.ugrid_meta(0, 0)
        self._counted = (0, 0, 0, 0)

    def run(self, tend: float, maxsteps: int) -> tuple[int, str]:
        err = N.err_buffer()
        status = self.lib.ugrid_run(self.handle, tend, maxsteps, C.byref(self.meta), err, N.ERR_LEN)
        return status, err.value.decode()

    def pull(self) -> None:
        s = self.solver
        err = N.err_buffer()
        u = np.empty_like(s.U)
        t = np.empty(s.ncase)
        dt = np.empty(s.ncase)
        if self.lib.ugrid_download(self.handle, N.dptr(u), u.size, N.dptr(t), N.dptr(dt), err, N.ERR_LEN):
            raise RuntimeError(err.value.decode())
        s.U[:] = u
        s.t[:] = t
        s.dt[:] = dt
        s.nsteps += self.meta.nsteps - self._counted[0]
        s.nreject += self.meta.nreject - self._counted[1]
        imp = self.implicit_stats()
        s.implicit["updates"] += imp["updates"] - self._counted[2]
        s.implicit["retries"] += imp["retries"] - self._counted[3]
        if imp["checks"]:
            for key in ("initial_defect", "final_defect", "tolerance"):
                s.implicit[key] = imp[key]
        self._counted = (self.meta.nsteps, self.meta.nreject, imp["updates"], imp["retries"])

    def implicit_stats(self) -> dict:
        info = N.ugrid_implicit_info()
        if self.lib.ugrid_implicit_stats(self.handle, C.byref(info)):
            raise RuntimeError("ugrid_implicit_stats failed")
        return {name: getattr(info, name) for name, _ in info._fields_}

    def rhs(self, u: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros_like(u)
        err = N.err_buffer()
        if self.lib.ugrid_rhs(self.handle, N.dptr(u), u.size, N.dptr(r), err, N.ERR_LEN):
            raise RuntimeError(err.value.decode())
        return r

    def launch_info(self) -> dict:
        info = N.ugrid_launch_info()
        self.lib.ugrid_get_launch(self.handle, C.byref(info))
        return dict(rows=info.rows, strips=info.strips, blocks=info.blocks, threads=info.threads)

    def close(self) -> None:
        if self.handle:
            self.lib.ugrid_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solver2d_cuda_available() -> tuple[bool, str]:
    """(available, reason) as solver2d_cuda_available (src/euler2d_cuda.cu:510-528)."""
    try:
        lib = N.load_library()
    except Exception as exc:  # library not built
        return False, str(exc)
    reason = N.err_buffer()
    ok = lib.ugrid_device_available(reason, N.ERR_LEN)
    return bool(ok), reason.value.decode()


def solver2d_run_cuda(s: Solver2D, tend: float, maxsteps: int) -> int:
    """Advance every case to tend on the device (src/euler2d_cuda.cu:530-708): uploads the state,
    runs, writes back U, t, dt, nsteps and nreject.  Refusals print the reference's message and
    return 3."""
    par = s.par
    if par.time_integrator not in (TIME_SSP_RK3, TIME_BACKWARD_EULER):
        s.last_error = "unknown time integrator"
    elif par.limiter != LIM_FIRST_ORDER:
        s.last_error = ("CUDA backend currently requires --limiter first; refusing to change the method "
                        "implicitly")
    else:
        ok, reason = solver2d_cuda_available()
        s.last_error = "" if ok else f"CUDA backend unavailable: {reason}"
        if ok and s.cut is not None and s.cut.nbad_nbhd != 0:
            s.last_error = f"CUDA backend refuses {s.cut.nbad_nbhd} invalid SRD neighborhoods"
    if s.last_error:
        print(s.last_error, file=sys.stderr)
        return 3
    try:
        session = DeviceSession(s)
    except RuntimeError as exc:  # a setup the device backend refuses (nothing ran)
        s.last_error = str(exc)
        print(s.last_error, file=sys.stderr)
        return 3
    with session:
        status, message = session.run(tend, maxsteps)
        # The reference writes the state, clocks and steps back after every device run, a step
        # that cannot advance included (src/euler2d_cuda.cu:694-707); refused controls stop first.
        if status in (0, 1, 2, 4) or (status == 3 and session.implicit_stats()["reason"] == 5):
            session.pull()
    s.last_error = message
    if status == 2 and par.time_integrator == TIME_SSP_RK3:
        print(f"CUDA solver: {message}", file=sys.stderr)
    elif status != 0 and status != 1 and message:
        print(message, file=sys.stderr)
    return status


def solver2d_run_backend(s: Solver2D, tend: float, maxsteps: int, backend: int) -> int:
    if backend == BACKEND_CUDA:
        return solver2d_run_cuda(s, tend, maxsteps)
    raise NotImplementedError("the CPU backend is the reference's own solver (src/euler2d.cpp:595-698)")
