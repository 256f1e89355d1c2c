"""Host mesh handles over the C ABI (amr_mesh_* in include/amr_hip.h).

`Mesh` owns one built mesh in host memory.  The same class also drives the reference's own
mesher through the test-only shim (prefix "ref_mesh"), which is how the tests show that the
two produce identical arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .config import Case


def params_of(case: Case) -> N.amr_params:
    """Time-step parameters of a case (the Params2D fields the path reads)."""
    p = N.amr_params()
    p.gamma = case.gamma
    p.cfl = case.cfl
    p.flux = case.flux
    p.limiter = 3
    p.ausm_mco = 1.0
    for k in range(4):
        p.bc[k] = 1  # legacy default (transmissive); every side sets its own kind
    p.dt_fixed = case.dt_fixed
    p.time_integrator = 0
    return p


SIDE_FIELDS = ("kind", "reference", "total_pressure", "total_temperature", "gas_constant", "direction",
               "outlet_pressure", "wall_velocity")


def boundaries_of(case: Case, sides: dict | None = None) -> "C.Array":
    """Side boundaries: freestream reference state and outlet pressure on all four sides
    (src/quick_airfoil.cpp:45-46).  `sides` maps a side (0 xmin, 1 xmax, 2 ymin, 3 ymax) to
    nondimensional field overrides, the settings the reference's static-AMR command line takes
    per side (src/static_amr2d_cli.cpp:149-170: kind, reference state, total pressure /
    temperature / gas constant / direction of a total inlet, outlet pressure, wall velocity)."""
    arr = (N.amr_boundary * 4)()
    ref = case.freestream
    for side in range(4):
        b = arr[side]
        b.kind = case.bc_kinds[side]
        for v in range(4):
            b.reference[v] = ref[v]
        b.total_pressure, b.total_temperature, b.gas_constant = 0.0, 0.0, 1.0
        b.direction[0], b.direction[1] = 1.0, 0.0
        b.outlet_pressure = case.outlet_p_star
        b.wall_velocity[0], b.wall_velocity[1] = 0.0, 0.0
    for side, fields in (sides or {}).items():
        b = arr[int(side)]
        for name, value in fields.items():
            if name not in SIDE_FIELDS:
                raise KeyError(f"unknown boundary field {name}")
            if name in ("reference", "direction", "wall_velocity"):
                target = getattr(b, name)
                for i, x in enumerate(value):
                    target[i] = float(x)
            else:
                setattr(b, name, int(value) if name == "kind" else float(value))
    return arr


def transport_of(case: Case) -> N.amr_transport:
    """Transport model (src/quick_airfoil.cpp:47-48)."""
    t = N.amr_transport()
    t.model = case.model
    t.law = case.law if case.model else 0
    t.stabilization = case.stabilization if case.model else 0
    t.mu_ref = case.mu_star if case.model else 0.0
    t.gas_constant = 1.0
    t.temperature_ref = 1.0 / case.gamma if case.law == 1 else 1.0
    t.sutherland = case.sutherland_star if case.law == 1 else 0.0
    t.viscous_cfl = case.viscous_cfl
    t.sa_model = case.sa_model
    t.sa_nu_inf = case.sa_nu_inf if case.sa_model else 0.0
    t.sa_cfl = case.sa_cfl
    t.wall_probe_min_h = case.wall_probe_min_h
    return t


def build_options(opts: dict):
    boxes = (N.amr_box * max(1, len(opts["boxes"])))()
    for k, b in enumerate(opts["boxes"]):
        boxes[k].x0, boxes[k].x1, boxes[k].y0, boxes[k].y1, boxes[k].level = b[0], b[1], b[2], b[3], int(b[4])
    o = N.amr_build_options()
    o.nx, o.ny, o.levels, o.buffer = opts["nx"], opts["ny"], opts["levels"], opts["buffer"]
    o.xmin, o.xmax, o.ymin, o.ymax = opts["xmin"], opts["xmax"], opts["ymin"], opts["ymax"]
    o.max_cells = opts["max_cells"]
    o.agglomerate_small = 1 if opts.get("agglomerate_small", False) else 0
    o.n_boxes = len(opts["boxes"])
    o.boxes = C.cast(boxes, C.POINTER(N.amr_box))
    return o, boxes


class MeshError(RuntimeError):
    pass


class Mesh:
    """One built cut-cell mesh (host memory, owned by the library)."""

    def __init__(self, lib, prefix, handle):
        self._lib, self._prefix, self.handle = lib, prefix, handle

    def _fn(self, name):
        return getattr(self._lib, f"{self._prefix}_{name}")

    @classmethod
    def build(cls, polygon: np.ndarray, patches: np.ndarray | None, options: dict, lib=None,
              prefix: str = "amr_mesh") -> "Mesh":
        lib = lib or N.load_library()
        poly = np.ascontiguousarray(polygon, dtype=np.float64)
        px, py = np.ascontiguousarray(poly[:, 0]), np.ascontiguousarray(poly[:, 1])
        pt = None if patches is None else np.ascontiguousarray(patches, dtype=np.int32)
        o, keep = build_options(options)
        handle = C.c_void_p()
        err = N.err_buffer()
        status = getattr(lib, f"{prefix}_build")(
            C.byref(o), N.dptr(px), N.dptr(py), None if pt is None else pt.ctypes.data_as(N.c_i32_p),
            len(px), C.byref(handle), err, N.ERR_LEN)
        del keep
        if status:
            raise MeshError(err.value.decode())
        return cls(lib, prefix, handle)

    def set_physics(self, case: Case, sides: dict | None = None):
        b, t = boundaries_of(case, sides), transport_of(case)
        if self._fn("set_physics")(self.handle, b, C.byref(t)):
            raise MeshError("set_physics failed")
        self.case = case
        return self

    def prepare_viscous(self, params: N.amr_params):
        err = N.err_buffer()
        if self._fn("prepare_viscous")(self.handle, C.byref(params), err, N.ERR_LEN):
            raise MeshError(err.value.decode())
        return self

    @property
    def view(self) -> N.amr_mesh_view:
        v = N.amr_mesh_view()
        if self._fn("get_view")(self.handle, C.byref(v)):
            raise MeshError("get_view failed")
        return v

    @property
    def info(self) -> N.amr_mesh_info:
        i = N.amr_mesh_info()
        self._fn("get_info")(self.handle, C.byref(i))
        return i

    @property
    def n_cells(self) -> int:
        return int(self.info.n_cells)

    def nv(self) -> int:
        return 5 if self.view.transport.sa_model else 4

    def initialize(self, params: N.amr_params, primitive) -> np.ndarray:
        u = np.zeros(self.nv() * self.n_cells)
        w = np.ascontiguousarray(primitive, dtype=np.float64)
        if self._fn("initialize")(self.handle, C.byref(params), N.dptr(w), N.dptr(u), u.size):
            raise MeshError("initialize failed")
        return u

    def arrays(self) -> dict:
        """Copies of every mesh and viscous-table array (numpy)."""
        v, i = self.view, self.info
        n, nf, nw = int(v.n_cells), int(v.n_faces), int(v.n_walls)
        out = {
            "cells": N.as_array(v.cells, n, N.CELL_DTYPE),
            "faces": N.as_array(v.faces, nf, N.FACE_DTYPE),
            "walls": N.as_array(v.walls, nw, N.WALL_DTYPE),
            "face_offsets": N.as_array(v.face_offsets, n + 1, np.uint64),
            "cell_faces": N.as_array(v.cell_faces, int(i.n_cell_faces), np.uint64),
            "group_offsets": N.as_array(v.group_offsets, n + 1, np.uint64),
            "group_members": N.as_array(v.group_members, int(i.n_group_members), np.uint64),
            "member_offsets": N.as_array(v.member_offsets, n + 1, np.uint64),
            "member_groups": N.as_array(v.member_groups, int(i.n_group_members), np.uint64),
            "group_volume": N.as_array(v.group_volume, n, np.float64),
        }
        if v.visc_offsets:
            out["visc_offsets"] = N.as_array(v.visc_offsets, n + 1, np.uint64)
            out["visc_terms"] = N.as_array(v.visc_terms, int(i.n_visc_terms), N.TERM_DTYPE)
            out["wall_samples"] = N.as_array(v.wall_samples, nw, N.SAMPLE_DTYPE)
            out["face_samples"] = N.as_array(v.face_samples, nf, N.SAMPLE_DTYPE)
            out["face_geometry"] = N.as_array(v.face_geometry, nf, N.GEOMETRY_DTYPE)
        if v.wall_distance:
            out["wall_distance"] = N.as_array(v.wall_distance, n, np.float64)
        part = np.zeros((int(i.n_partition), 5), dtype=np.int32)
        self._fn("copy_partition")(self.handle, part.ctypes.data_as(N.c_i32_p), part.shape[0])
        leaf = np.zeros((int(i.n_leaf_components), 2), dtype=np.uint64)
        self._fn("copy_leaf_components")(self.handle, leaf.ctypes.data_as(N.c_u64_p), leaf.shape[0])
        agg = np.zeros((int(i.n_agglomeration), 9), dtype=np.float64)
        if agg.size:
            self._fn("copy_agglomeration")(self.handle, N.dptr(agg), agg.shape[0])
        out.update(partition=part, leaf_components=leaf, agglomeration=agg)
        return out

    def close(self):
        if self.handle:
            self._fn("destroy")(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_case_mesh(case: Case, polygon: np.ndarray, patches: np.ndarray | None, lib=None,
                    prefix: str = "amr_mesh", sides: dict | None = None) -> Mesh:
    """Mesh + physics + viscous tables for one case, as the solver executable sets them up
    (`sides`: per-side boundary overrides, see boundaries_of)."""
    mesh = Mesh.build(polygon, patches, case.mesh_options, lib=lib, prefix=prefix)
    mesh.set_physics(case, sides)
    mesh.prepare_viscous(params_of(case))
    return mesh
