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
