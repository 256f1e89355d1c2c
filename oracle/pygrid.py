"""TEST INFRASTRUCTURE ONLY: ctypes bindings of the tensor-product grid oracle (grid_oracle.cpp in
liboracle.so) and of the reference's legacy grid solver built from its sources
(_ref/libamr_ref.so, ref_grid_* entry points; build container only).

Problems are described by the package's euler2d.Solver2D (grid, parameters, per-case arrays),
used here only as a data container.  Only tests/ import this module, as the checker.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import _native as N

from . import pyoracle

_bound = set()


def _oracle():
    lib = pyoracle.oracle_lib()
    if "oracle" not in _bound:
        S = C.POINTER(N.ugrid_setup)
        lib.grid_oracle_apply_bc.restype = None
        lib.grid_oracle_apply_bc.argtypes = [S, N.c_double_p]
        lib.grid_oracle_rhs.restype = C.c_int
        lib.grid_oracle_rhs.argtypes = [S, N.c_double_p, N.c_double_p]
        lib.grid_oracle_run.restype = C.c_int
        lib.grid_oracle_run.argtypes = [S, N.c_double_p, N.c_double_p, N.c_double_p, C.c_double, C.c_int64,
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        lib.grid_oracle_run_implicit.restype = C.c_int
        lib.grid_oracle_run_implicit.argtypes = [S, N.c_double_p, N.c_double_p, N.c_double_p, C.c_double, C.c_int64,
                                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64), N.c_double_p]
        lib.grid_oracle_flux.restype = None
        lib.grid_oracle_flux.argtypes = [C.c_int, N.c_double_p, N.c_double_p, C.c_double, C.c_double, N.c_double_p]
        _bound.add("oracle")
    return lib


def _ref():
    lib = pyoracle.ref_lib()
    if lib is None:
        return None
    if "ref" not in _bound:
        P, I, D = C.POINTER(N.ugrid_params), C.POINTER(C.c_int), N.c_double_p
        lib.ref_grid_geometry.restype = C.c_int
        lib.ref_grid_geometry.argtypes = [C.c_int, C.c_int, D, P, I, D, D, D, D, D, D, D]
        lib.ref_grid_apply_bc.restype = None
        lib.ref_grid_apply_bc.argtypes = [C.c_int, C.c_int, C.c_int, P, D]
        lib.ref_grid_rhs.restype = C.c_int
        lib.ref_grid_rhs.argtypes = [C.c_int, C.c_int, D, C.c_int, P, I, D, D, D, D]
        lib.ref_grid_run.restype = C.c_int
        lib.ref_grid_run.argtypes = [C.c_int, C.c_int, D, C.c_int, P, I, D, D, D, D, D, C.c_double, C.c_int64,
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        lib.ref_grid_run_implicit.restype = C.c_int
        lib.ref_grid_run_implicit.argtypes = [C.c_int, C.c_int, D, C.c_int, P, I, D, D, D, D, D, C.c_double,
                                              C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64), D]
        lib.ref_grid_cut_build.restype = C.c_int
        lib.ref_grid_cut_build.argtypes = [C.c_int, C.c_int, D, P, I, D, D, D, C.c_int, C.POINTER(C.c_int),
                                           C.POINTER(D), C.POINTER(D), C.POINTER(D), C.POINTER(D), D, C.c_int,
                                           C.POINTER(C.c_int64), C.POINTER(N.amr_wall), C.c_int]
        lib.ref_grid_cut_rhs.restype = C.c_int
        lib.ref_grid_cut_rhs.argtypes = [C.c_int, C.c_int, D, C.c_int, P, I, D, D, D, D, C.c_int, D, D]
        lib.ref_grid_cut_run.restype = C.c_int
        lib.ref_grid_cut_run.argtypes = [C.c_int, C.c_int, D, C.c_int, P, I, D, D, D, D, C.c_int, D, D, D,
                                         C.c_double, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int64), D]
        lib.ref_flux.restype = None
        lib.ref_flux.argtypes = [C.c_int, D, D, C.c_double, C.c_double, D]
        _bound.add("ref")
    return lib


def ref_available() -> bool:
    return pyoracle.ref_available()


# ---------------------------------------------------------------- the oracle --
def apply_bc(s, u: np.ndarray) -> np.ndarray:
    u = np.array(u, dtype=np.float64, copy=True)
    setup, keep = s._setup()
    _oracle().grid_oracle_apply_bc(C.byref(setup), N.dptr(u))
    return u


def rhs(s, u: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=np.float64)
    r = np.zeros_like(u)
    setup, keep = s._setup()
    _oracle().grid_oracle_rhs(C.byref(setup), N.dptr(u), N.dptr(r))
    return r


def run(s, tend: float, maxsteps: int) -> dict:
    """The oracle's run from s's current U, t, dt (s is not modified)."""
    u = np.array(s.U, dtype=np.float64, copy=True)
    t = np.array(s.t, dtype=np.float64, copy=True)
    dt = np.array(s.dt, dtype=np.float64, copy=True)
    nsteps, nreject = C.c_int64(0), C.c_int64(0)
    setup, keep = s._setup()
    status = _oracle().grid_oracle_run(C.byref(setup), N.dptr(u), N.dptr(t), N.dptr(dt), tend, maxsteps,
                                       C.byref(nsteps), C.byref(nreject))
    return dict(u=u, t=t, dt=dt, status=status, nsteps=nsteps.value, nreject=nreject.value)


def run_implicit(s, tend: float, maxsteps: int) -> dict:
    """The oracle's backward-Euler run from s's current U, t, dt (s is not modified)."""
    u = np.array(s.U, dtype=np.float64, copy=True)
    t = np.array(s.t, dtype=np.float64, copy=True)
    dt = np.array(s.dt, dtype=np.float64, copy=True)
    nsteps = C.c_int64(0)
    counts = (C.c_int64 * 4)(0, 0, 0, 0)
    defects = np.zeros(3)
    setup, keep = s._setup()
    status = _oracle().grid_oracle_run_implicit(C.byref(setup), N.dptr(u), N.dptr(t), N.dptr(dt), tend, maxsteps,
                                                C.byref(nsteps), counts, N.dptr(defects))
    return dict(u=u, t=t, dt=dt, status=status, nsteps=nsteps.value, updates=counts[0], retries=counts[1],
                checks=counts[2], reason=counts[3], initial_defect=defects[0], final_defect=defects[1],
                tolerance=defects[2])


def flux(kind: int, left, right, gamma: float = 1.4, mco: float = 1.0) -> np.ndarray:
    a = np.ascontiguousarray(left, dtype=np.float64)
    b = np.ascontiguousarray(right, dtype=np.float64)
    f = np.zeros(4)
    _oracle().grid_oracle_flux(kind, N.dptr(a), N.dptr(b), gamma, mco, N.dptr(f))
    return f


# ------------------------------------------------------------- the reference --
def _stretch(s):
    types = (C.c_int * 2)(s.par.sx.type, s.par.sy.type)
    params = np.array([s.par.sx.param, s.par.sy.param], dtype=np.float64)
    return types, params


def _box(s) -> np.ndarray:
    g = s.grid
    return np.array([g.xmin, g.xmax, g.ymin, g.ymax], dtype=np.float64)


def ref_geometry(s) -> dict:
    g = s.grid
    out = {k: np.zeros(n) for k, n in (("xf", g.nx + 1), ("xc", g.mx), ("dx", g.mx), ("yf", g.ny + 1),
                                          ("yc", g.my), ("dy", g.my))}
    types, params = _stretch(s)
    box = _box(s)
    p = s.par.to_c()
    _ref().ref_grid_geometry(g.nx, g.ny, N.dptr(box), C.byref(p), types, N.dptr(params),
                             *(N.dptr(out[k]) for k in ("xf", "xc", "dx", "yf", "yc", "dy")))
    return out


def ref_apply_bc(s, u: np.ndarray) -> np.ndarray:
    u = np.array(u, dtype=np.float64, copy=True)
    p = s.par.to_c()
    _ref().ref_grid_apply_bc(s.grid.nx, s.grid.ny, s.ncase, C.byref(p), N.dptr(u))
    return u


def ref_rhs(s, u: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=np.float64)
    r = np.zeros_like(u)
    types, params = _stretch(s)
    box = _box(s)
    p = s.par.to_c()
    mco = np.ascontiguousarray(s.mco, dtype=np.float64)
    _ref().ref_grid_rhs(s.grid.nx, s.grid.ny, N.dptr(box), s.ncase, C.byref(p), types, N.dptr(params), N.dptr(mco),
                        N.dptr(u), N.dptr(r))
    return r


def ref_run(s, tend: float, maxsteps: int) -> dict:
    """The reference's CPU solver2d_run from s's current state (s is not modified)."""
    u = np.array(s.U, dtype=np.float64, copy=True)
    t = np.array(s.t, dtype=np.float64, copy=True)
    dt = np.array(s.dt, dtype=np.float64, copy=True)
    types, params = _stretch(s)
    box = _box(s)
    p = s.par.to_c()
    mco = np.ascontiguousarray(s.mco, dtype=np.float64)
    nsteps, nreject = C.c_int64(0), C.c_int64(0)
    status = _ref().ref_grid_run(s.grid.nx, s.grid.ny, N.dptr(box), s.ncase, C.byref(p), types, N.dptr(params),
                                 N.dptr(mco), N.dptr(u), N.dptr(t), N.dptr(dt), tend, maxsteps, C.byref(nsteps),
                                 C.byref(nreject))
    return dict(u=u, t=t, dt=dt, status=status, nsteps=nsteps.value, nreject=nreject.value)


def ref_run_implicit(s, tend: float, maxsteps: int) -> dict:
    """The reference's CPU backward-Euler driver from s's current state (s is not modified)."""
    u = np.array(s.U, dtype=np.float64, copy=True)
    t = np.array(s.t, dtype=np.float64, copy=True)
    dt = np.array(s.dt, dtype=np.float64, copy=True)
    types, params = _stretch(s)
    box = _box(s)
    p = s.par.to_c()
    mco = np.ascontiguousarray(s.mco, dtype=np.float64)
    nsteps = C.c_int64(0)
    counts = (C.c_int64 * 2)(0, 0)
    defects = np.zeros(3)
    status = _ref().ref_grid_run_implicit(s.grid.nx, s.grid.ny, N.dptr(box), s.ncase, C.byref(p), types,
                                          N.dptr(params), N.dptr(mco), N.dptr(u), N.dptr(t), N.dptr(dt), tend,
                                          maxsteps, C.byref(nsteps), counts, N.dptr(defects))
    return dict(u=u, t=t, dt=dt, status=status, nsteps=nsteps.value, updates=counts[0], retries=counts[1],
                initial_defect=defects[0], final_defect=defects[1], tolerance=defects[2])


def ref_flux(kind: int, left, right, gamma: float = 1.4, mco: float = 1.0) -> np.ndarray:
    a = np.ascontiguousarray(left, dtype=np.float64)
    b = np.ascontiguousarray(right, dtype=np.float64)
    f = np.zeros(4)
    _ref().ref_flux(kind, N.dptr(a), N.dptr(b), gamma, mco, N.dptr(f))
    return f


def ref_cut_build(s, px, py, patches=None, capacity: int = 4096, capacity_walls: int = 8192) -> dict:
    """The reference's cut2d_build_polygon on s's grid: every array (the counts under 'counts')."""
    g = s.grid
    cells, nfx, nfy = g.mx * g.my, (g.nx + 1) * g.ny, g.nx * (g.ny + 1)
    cell = [np.zeros(cells) for _ in range(10)]
    xface = [np.zeros(nfx) for _ in range(2)]
    yface = [np.zeros(nfy) for _ in range(2)]
    small = [np.zeros(capacity) for _ in range(5)]
    smem = np.zeros(25 * capacity)
    counts = (C.c_int64 * 5)()
    walls = (N.amr_wall * capacity_walls)()
    ptrs = lambda arrays: (N.c_double_p * len(arrays))(*[N.dptr(a) for a in arrays])
    types, params = _stretch(s)
    box = _box(s)
    p = s.par.to_c()
    px = np.ascontiguousarray(px, dtype=np.float64)
    py = np.ascontiguousarray(py, dtype=np.float64)
    pat = None if patches is None else (C.c_int * len(patches))(*patches)
    status = _ref().ref_grid_cut_build(g.nx, g.ny, N.dptr(box), C.byref(p), types, N.dptr(params), N.dptr(px),
                                       N.dptr(py), len(px), pat, ptrs(cell), ptrs(xface), ptrs(yface), ptrs(small),
                                       N.dptr(smem), capacity, counts, walls, capacity_walls)
    out = dict(status=status)
    if status:
        return out
    ns = counts[0]
    out.update(zip(("vol", "cx", "cy", "flag", "wax", "way", "wcx", "wcy", "nbrcount", "smallpos"), cell))
    out.update(axl=xface[0], axy=xface[1], ayl=yface[0], ayx=yface[1], smem=smem[:25 * ns])
    out.update(zip(("sidx", "scnt", "svhat", "sxhat", "syhat"), [a[:ns] for a in small]))
    out.update(nsmall=ns, nbad_nbhd=counts[1], nsolid=counts[2], ncut=counts[3])
    out["walls"] = N.as_array(C.cast(walls, C.POINTER(N.amr_wall)), counts[4], N.WALL_DTYPE)
    return out


def ref_cut_rhs(s, px, py, u: np.ndarray) -> np.ndarray:
    """The reference's residual with the cut geometry of the polygon (built by the reference)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    r = np.zeros_like(u)
    types, params = _stretch(s)
    box = _box(s)
    p = s.par.to_c()
    mco = np.ascontiguousarray(s.mco, dtype=np.float64)
    px = np.ascontiguousarray(px, dtype=np.float64)
    py = np.ascontiguousarray(py, dtype=np.float64)
    assert _ref().ref_grid_cut_rhs(s.grid.nx, s.grid.ny, N.dptr(box), s.ncase, C.byref(p), types, N.dptr(params),
                                   N.dptr(mco), N.dptr(px), N.dptr(py), len(px), N.dptr(u), N.dptr(r)) == 0
    return r


def ref_cut_run(s, px, py, tend: float, maxsteps: int) -> dict:
    """The reference's CPU solver2d_run (RK or backward Euler) with the polygon's cut geometry."""
    u = np.array(s.U, dtype=np.float64, copy=True)
    t = np.array(s.t, dtype=np.float64, copy=True)
    dt = np.array(s.dt, dtype=np.float64, copy=True)
    types, params = _stretch(s)
    box = _box(s)
    p = s.par.to_c()
    mco = np.ascontiguousarray(s.mco, dtype=np.float64)
    px = np.ascontiguousarray(px, dtype=np.float64)
    py = np.ascontiguousarray(py, dtype=np.float64)
    nsteps, nreject = C.c_int64(0), C.c_int64(0)
    counts = (C.c_int64 * 2)(0, 0)
    defects = np.zeros(3)
    status = _ref().ref_grid_cut_run(s.grid.nx, s.grid.ny, N.dptr(box), s.ncase, C.byref(p), types, N.dptr(params),
                                     N.dptr(mco), N.dptr(px), N.dptr(py), len(px), N.dptr(u), N.dptr(t), N.dptr(dt),
                                     tend, maxsteps, C.byref(nsteps), C.byref(nreject), counts, N.dptr(defects))
    return dict(u=u, t=t, dt=dt, status=status, nsteps=nsteps.value, nreject=nreject.value, updates=counts[0],
                retries=counts[1], initial_defect=defects[0], final_defect=defects[1], tolerance=defects[2])
