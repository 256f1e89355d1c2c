"""ctypes mirror of include/amr_hip.h and the loader of the in-tree libamr_hip.so.

The structures below are byte-for-byte the C records of amr_hip.h (which are themselves
layout-identical to the reference's C++ structs, src/static_amr2d.h:37-45 and friends).
numpy structured dtypes of the same records let tests read mesh arrays without copies.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PACKAGE_DIR = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE_DIR / "libamr_hip.so"

c_double_p = C.POINTER(C.c_double)
c_u64_p = C.POINTER(C.c_uint64)
c_i32_p = C.POINTER(C.c_int32)


class amr_cell(C.Structure):
    _fields_ = [("ix", C.c_int32), ("iy", C.c_int32), ("span", C.c_int32), ("level", C.c_int32),
                ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double), ("y1", C.c_double),
                ("volume", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("wall_begin", C.c_uint64), ("wall_count", C.c_uint64)]


class amr_face(C.Structure):
    _fields_ = [("left", C.c_int32), ("right", C.c_int32), ("axis", C.c_int32), ("boundary", C.c_int32),
                ("coordinate", C.c_double), ("lo", C.c_double), ("hi", C.c_double),
                ("length", C.c_double), ("centroid", C.c_double)]


class amr_wall(C.Structure):
    _fields_ = [("owner_cell", C.c_uint64), ("source_edge_id", C.c_int32), ("body_id", C.c_int32),
                ("patch_id", C.c_int32), ("x0", C.c_double), ("y0", C.c_double), ("x1", C.c_double),
                ("y1", C.c_double), ("nx", C.c_double), ("ny", C.c_double), ("length", C.c_double)]


class amr_grad_term(C.Structure):
    _fields_ = [("source", C.c_int32), ("side", C.c_int32), ("x", C.c_double), ("y", C.c_double)]


class amr_wall_sample(C.Structure):
    _fields_ = [("donor", C.c_int32), ("distance", C.c_double)]


class amr_face_geometry(C.Structure):
    _fields_ = [("dx", C.c_double), ("dy", C.c_double), ("ex", C.c_double), ("ey", C.c_double),
                ("distance", C.c_double)]


class amr_boundary(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reference", C.c_double * 4), ("total_pressure", C.c_double),
                ("total_temperature", C.c_double), ("gas_constant", C.c_double),
                ("direction", C.c_double * 2), ("outlet_pressure", C.c_double),
                ("wall_velocity", C.c_double * 2)]


class amr_transport(C.Structure):
    _fields_ = [("model", C.c_int32), ("law", C.c_int32), ("stabilization", C.c_int32),
                ("mu_ref", C.c_double), ("gas_constant", C.c_double), ("temperature_ref", C.c_double),
                ("sutherland", C.c_double), ("viscous_cfl", C.c_double), ("sa_model", C.c_int32),
                ("sa_nu_inf", C.c_double), ("sa_cfl", C.c_double), ("wall_probe_min_h", C.c_double)]


class amr_params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("flux", C.c_int32), ("limiter", C.c_int32),
                ("ausm_mco", C.c_double), ("bc", C.c_int32 * 4), ("dt_fixed", C.c_double),
                ("time_integrator", C.c_int32)]


class amr_mesh_view(C.Structure):
    _fields_ = [("n_cells", C.c_uint64), ("n_faces", C.c_uint64), ("n_walls", C.c_uint64),
                ("cells", C.POINTER(amr_cell)), ("faces", C.POINTER(amr_face)), ("walls", C.POINTER(amr_wall)),
                ("face_offsets", c_u64_p), ("cell_faces", c_u64_p),
                ("group_offsets", c_u64_p), ("group_members", c_u64_p),
                ("member_offsets", c_u64_p), ("member_groups", c_u64_p),
                ("group_volume", c_double_p), ("small_cells", C.c_uint64), ("agglomerate_small", C.c_int32),
                ("boundaries", amr_boundary * 4), ("transport", amr_transport),
                ("visc_offsets", c_u64_p), ("visc_terms", C.POINTER(amr_grad_term)),
                ("wall_samples", C.POINTER(amr_wall_sample)), ("face_samples", C.POINTER(amr_wall_sample)),
                ("face_geometry", C.POINTER(amr_face_geometry)), ("wall_distance", c_double_p)]


class amr_box(C.Structure):
    _fields_ = [("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double), ("y1", C.c_double),
                ("level", C.c_int32)]


class amr_build_options(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("levels", C.c_int32), ("buffer", C.c_int32),
                ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
                ("max_cells", C.c_uint64), ("agglomerate_small", C.c_int32), ("n_boxes", C.c_int32),
                ("boxes", C.POINTER(amr_box))]


class amr_mesh_info(C.Structure):
    _fields_ = [(name, C.c_uint64) for name in (
        "n_cells", "n_faces", "n_walls", "n_cell_faces", "n_group_members", "n_visc_terms",
        "background_leaves", "split_leaves", "coarse_fine_faces", "small_cells",
        "n_partition", "n_leaf_components", "n_agglomeration")]


class amr_session_options(C.Structure):
    _fields_ = [("batch_trials", C.c_int32), ("graph", C.c_int32), ("event_timing", C.c_int32),
                ("precision", C.c_int32)]


class amr_state_meta(C.Structure):
    _fields_ = [("time", C.c_double), ("steps", C.c_int64), ("rejected", C.c_int64)]


class amr_gpu_stats(C.Structure):
    _fields_ = [("allocated_bytes", C.c_uint64), ("pinned_bytes", C.c_uint64),
                ("geometry_h2d_bytes", C.c_uint64), ("state_h2d_bytes", C.c_uint64),
                ("control_h2d_bytes", C.c_uint64), ("control_d2h_bytes", C.c_uint64),
                ("state_d2h_bytes", C.c_uint64), ("status_polls", C.c_uint64), ("full_downloads", C.c_uint64),
                ("setup_seconds", C.c_double), ("advance_seconds", C.c_double),
                ("download_seconds", C.c_double), ("advance_event_seconds", C.c_double),
                ("event_batches", C.c_uint64), ("graph_nodes_per_trial", C.c_uint64)]


class amr_steady_options(C.Structure):
    _fields_ = [("mach", C.c_double), ("aoa_degrees", C.c_double), ("moment", C.c_double * 2),
                ("region", C.c_double * 4), ("raw_tol", C.c_double), ("load_tol", C.c_double),
                ("min_time", C.c_double), ("period", C.c_double), ("hold", C.c_int32), ("mode", C.c_int32),
                ("stop_time", C.c_double), ("budget_seconds", C.c_double)]


class amr_steady_check(C.Structure):
    _fields_ = [("time", C.c_double), ("steps", C.c_int64), ("rejected", C.c_int64), ("elapsed", C.c_double),
                ("residual", C.c_double * 5), ("cl", C.c_double), ("cd", C.c_double), ("cm", C.c_double),
                ("dcl", C.c_double), ("dcd", C.c_double), ("dcm", C.c_double), ("hold", C.c_int32),
                ("pad", C.c_int32)]


class amr_steady_result(C.Structure):
    _fields_ = [("converged", C.c_int32), ("failed", C.c_int32), ("reason", C.c_int32), ("hold", C.c_int32),
                ("checks", C.c_int32), ("pad", C.c_int32), ("time", C.c_double), ("steps", C.c_int64),
                ("rejected", C.c_int64), ("residual", C.c_double * 5), ("cl", C.c_double), ("cd", C.c_double),
                ("cm", C.c_double), ("elapsed_seconds", C.c_double)]


class amr_kernel_time(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("total_ms", C.c_double), ("launches", C.c_int64)]


class amr_slot_capacity(C.Structure):
    _fields_ = [("cells", C.c_int32), ("faces", C.c_int32), ("walls", C.c_int32), ("csr", C.c_int32),
                ("terms", C.c_int32), ("groups", C.c_int32), ("members", C.c_int32), ("pad", C.c_int32)]


class amr_batch_features(C.Structure):
    _fields_ = [("laminar", C.c_int32), ("sa", C.c_int32), ("small", C.c_int32), ("visc_group", C.c_int32)]


class amr_batch_options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("tile", C.c_int32), ("graph", C.c_int32), ("batch_trials", C.c_int32)]


class amr_slot_state(C.Structure):
    _fields_ = [("time", C.c_double), ("target", C.c_double), ("steps", C.c_int64), ("rejected", C.c_int64),
                ("n_cells", C.c_int32), ("finished", C.c_int32), ("fatal", C.c_int32), ("pad", C.c_int32)]


STOP_REASONS = {0: "raw-residual-and-three-coefficients", 1: "time-budget", 2: "physical-time-cap",
                3: "admissibility-or-advance-failure", 4: "short-validation"}


def _dtype_of(struct) -> np.dtype:
    """numpy structured dtype with the exact offsets of a ctypes Structure."""
    names, formats, offsets = [], [], []
    for name, ctype in struct._fields_:
        names.append(name)
        formats.append(np.dtype(ctype))
        offsets.append(getattr(struct, name).offset)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": C.sizeof(struct)})


CELL_DTYPE = _dtype_of(amr_cell)
FACE_DTYPE = _dtype_of(amr_face)
WALL_DTYPE = _dtype_of(amr_wall)
TERM_DTYPE = _dtype_of(amr_grad_term)
SAMPLE_DTYPE = _dtype_of(amr_wall_sample)
GEOMETRY_DTYPE = _dtype_of(amr_face_geometry)

ERR_LEN = 512


def err_buffer():
    return C.create_string_buffer(ERR_LEN)


def as_array(pointer, count, dtype) -> np.ndarray:
    """Copy `count` records behind a ctypes pointer into a numpy array (empty when NULL)."""
    dtype = np.dtype(dtype)
    if count == 0 or not pointer:
        return np.zeros(0, dtype=dtype)
    raw = C.string_at(C.cast(pointer, C.c_void_p), int(count) * dtype.itemsize)
    return np.frombuffer(raw, dtype=dtype).copy()


def _bind_mesh_api(lib, prefix, handle_name):
    """Declare the amr_mesh_* style entry points (also used for the reference shim)."""
    H = C.c_void_p
    sig = {
        "build": (C.c_int, [C.POINTER(amr_build_options), c_double_p, c_double_p, c_i32_p, C.c_int32,
                            C.POINTER(H), C.c_char_p, C.c_size_t]),
        "set_physics": (C.c_int, [H, C.POINTER(amr_boundary), C.POINTER(amr_transport)]),
        "prepare_viscous": (C.c_int, [H, C.POINTER(amr_params), C.c_char_p, C.c_size_t]),
        "get_view": (C.c_int, [H, C.POINTER(amr_mesh_view)]),
        "get_info": (C.c_int, [H, C.POINTER(amr_mesh_info)]),
        "initialize": (C.c_int, [H, C.POINTER(amr_params), c_double_p, c_double_p, C.c_uint64]),
        "copy_partition": (C.c_int, [H, c_i32_p, C.c_uint64]),
        "copy_leaf_components": (C.c_int, [H, c_u64_p, C.c_uint64]),
        "copy_agglomeration": (C.c_int, [H, c_double_p, C.c_uint64]),
        "destroy": (None, [H]),
    }
    for name, (restype, argtypes) in sig.items():
        fn = getattr(lib, f"{prefix}_{name}")
        fn.restype = restype
        fn.argtypes = argtypes


def _bind_batch_api(lib):
    """Declare the amr_batch_* and amr_driver_* entry points."""
    H = C.c_void_p
    E = [C.c_char_p, C.c_size_t]
    sig = {
        "amr_batch_capacity": (C.c_int, [C.POINTER(amr_mesh_view), C.c_int32, C.POINTER(amr_slot_capacity),
                                         C.POINTER(amr_batch_features)]),
        "amr_batch_bytes": (C.c_uint64, [C.POINTER(amr_slot_capacity), C.c_int32, C.POINTER(amr_batch_features),
                                         C.c_int32, C.c_int32]),
        "amr_batch_create": (C.c_int, [C.POINTER(amr_mesh_view), C.POINTER(amr_params), C.POINTER(c_double_p),
                                       C.POINTER(amr_state_meta), C.c_int32, C.POINTER(amr_batch_options),
                                       C.POINTER(H), *E]),
        "amr_batch_create_slots": (C.c_int, [C.POINTER(amr_slot_capacity), C.c_int32, C.POINTER(amr_batch_features),
                                             C.POINTER(amr_batch_options), C.POINTER(H), *E]),
        "amr_batch_admit": (C.c_int, [H, C.c_int32, C.POINTER(amr_mesh_view), C.POINTER(amr_params), c_double_p,
                                      C.c_uint64, C.POINTER(amr_state_meta), *E]),
        "amr_batch_release": (C.c_int, [H, C.c_int32, *E]),
        "amr_batch_set_targets": (C.c_int, [H, c_double_p, C.POINTER(C.c_int64), *E]),
        "amr_batch_advance": (C.c_int, [H, C.c_int32, C.c_double, C.POINTER(amr_slot_state), *E]),
        "amr_batch_run": (C.c_int, [H, C.c_int32, C.POINTER(amr_slot_state), *E]),
        "amr_batch_set_launch_shape": (C.c_int, [H, C.c_int32, C.c_int32, C.c_int32, *E]),
        "amr_batch_slots": (C.c_int, [H, C.POINTER(amr_slot_state)]),
        "amr_batch_size": (C.c_int32, [H]),
        "amr_batch_download": (C.c_int, [H, C.c_int32, c_double_p, C.c_uint64, C.POINTER(amr_state_meta), *E]),
        "amr_batch_monitor": (C.c_int, [H, c_i32_p, c_double_p, c_double_p, c_double_p, c_double_p, c_i32_p, *E]),
        "amr_batch_profile": (C.c_int, [H, C.c_int32, C.POINTER(amr_kernel_time), C.c_int32, C.POINTER(C.c_int32),
                                        *E]),
        "amr_batch_stats": (None, [H, C.POINTER(amr_gpu_stats)]),
        "amr_batch_destroy": (None, [H]),
        "amr_driver_create": (C.c_int, [H, C.POINTER(H), *E]),
        "amr_driver_begin": (C.c_int, [H, C.c_int32, C.POINTER(amr_steady_options), *E]),
        "amr_driver_poll": (C.c_int, [H, C.c_double, C.POINTER(amr_steady_result), c_i32_p, C.POINTER(C.c_int32),
                                      *E]),
        "amr_driver_running": (C.c_int32, [H]),
        "amr_driver_destroy": (None, [H]),
        "amr_batch_solve": (C.c_int, [H, C.POINTER(amr_steady_options), C.POINTER(amr_steady_result), *E]),
    }
    for name, (restype, argtypes) in sig.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes


class ugrid_params(C.Structure):
    """ugrid_hip.h: the solver fields of the reference's Params2D (src/euler2d.h:38-52)."""
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("flux", C.c_int32), ("limiter", C.c_int32),
                ("ausm_mco", C.c_double), ("bc", C.c_int32 * 4), ("dt_fixed", C.c_double),
                ("time_integrator", C.c_int32), ("implicit_maxiter", C.c_int32), ("pseudo_cfl", C.c_double),
                ("implicit_rtol", C.c_double), ("implicit_atol", C.c_double)]


class ugrid_cut_view(C.Structure):
    _fields_ = [("nsmall", C.c_int32), ("nbad_nbhd", C.c_int32), ("nsolid", C.c_int32), ("ncut", C.c_int32),
                ("n_walls", C.c_uint64)] + [(name, c_double_p) for name in (
                    "vol", "cx", "cy", "flag", "axl", "axy", "ayl", "ayx", "wax", "way", "wcx", "wcy", "nbrcount",
                    "smallpos", "sidx", "scnt", "smem", "svhat", "sxhat", "syhat")] + [("walls", C.POINTER(amr_wall))]


class ugrid_setup(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("ncase", C.c_int32), ("pad", C.c_int32),
                ("dx", c_double_p), ("dy", c_double_p), ("params", ugrid_params), ("mco", c_double_p),
                ("cut", C.POINTER(ugrid_cut_view))]


class ugrid_axes(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("xf", c_double_p), ("xc", c_double_p), ("dx", c_double_p),
                ("yf", c_double_p), ("yc", c_double_p), ("dy", c_double_p)]


class ugrid_meta(C.Structure):
    _fields_ = [("nsteps", C.c_int64), ("nreject", C.c_int64)]


class ugrid_implicit_info(C.Structure):
    _fields_ = [("updates", C.c_int64), ("retries", C.c_int64), ("checks", C.c_int64), ("reason", C.c_int64),
                ("initial_defect", C.c_double), ("final_defect", C.c_double), ("tolerance", C.c_double)]


class ugrid_launch_info(C.Structure):
    _fields_ = [("rows", C.c_int32), ("strips", C.c_int32), ("blocks", C.c_int32), ("threads", C.c_int32)]


def _bind_ugrid_api(lib):
    H = C.c_void_p
    sig = {
        "ugrid_axis": (C.c_int, [C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_double, C.c_int32, C.c_int32,
                                 c_double_p, c_double_p, c_double_p, C.c_char_p, C.c_size_t]),
        "ugrid_create": (C.c_int, [C.POINTER(ugrid_setup), c_double_p, C.c_uint64, c_double_p, c_double_p,
                                   C.POINTER(H), C.c_char_p, C.c_size_t]),
        "ugrid_run": (C.c_int, [H, C.c_double, C.c_int64, C.POINTER(ugrid_meta), C.c_char_p, C.c_size_t]),
        "ugrid_download": (C.c_int, [H, c_double_p, C.c_uint64, c_double_p, c_double_p, C.c_char_p, C.c_size_t]),
        "ugrid_rhs": (C.c_int, [H, c_double_p, C.c_uint64, c_double_p, C.c_char_p, C.c_size_t]),
        "ugrid_destroy": (None, [H]),
        "ugrid_get_launch": (C.c_int, [H, C.POINTER(ugrid_launch_info)]),
        "ugrid_device_available": (C.c_int, [C.c_char_p, C.c_size_t]),
        "ugrid_profile": (C.c_int, [H, C.c_int32, c_double_p, C.c_char_p, C.c_size_t]),
        "ugrid_implicit_stats": (C.c_int, [H, C.POINTER(ugrid_implicit_info)]),
        "ugrid_cut_build": (C.c_int, [C.POINTER(ugrid_axes), c_double_p, c_double_p, C.c_int32, c_i32_p,
                                      C.POINTER(H), C.c_char_p, C.c_size_t]),
        "ugrid_cut_get": (C.c_int, [H, C.POINTER(ugrid_cut_view)]),
        "ugrid_cut_destroy": (None, [H]),
        "ugrid_solver2d_layout": (C.c_int, [C.POINTER(C.c_int64), C.c_int32]),
    }
    for name, (restype, argtypes) in sig.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes


_LIB = None


def load_library(path: os.PathLike | None = None):
    """Load libamr_hip.so (built in-tree by __graft_entry__.build()); raise if it is missing."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    target = Path(path) if path else LIBRARY_PATH
    if not target.exists():
        raise RuntimeError(f"{target} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(str(target))
    _bind_mesh_api(lib, "amr_mesh", "amr_mesh")
    H = C.c_void_p
    lib.amr_hip_create.restype = C.c_int
    lib.amr_hip_create.argtypes = [C.POINTER(amr_mesh_view), C.POINTER(amr_params), c_double_p, C.c_uint64,
                                   C.POINTER(amr_state_meta), C.POINTER(amr_session_options), C.POINTER(H),
                                   C.c_char_p, C.c_size_t]
    lib.amr_hip_advance.restype = C.c_int
    lib.amr_hip_advance.argtypes = [H, C.c_double, C.c_int64, C.c_double, C.POINTER(amr_state_meta),
                                    C.c_char_p, C.c_size_t]
    lib.amr_hip_download.restype = C.c_int
    lib.amr_hip_download.argtypes = [H, c_double_p, C.c_uint64, C.POINTER(amr_state_meta), C.c_char_p,
                                     C.c_size_t]
    lib.amr_hip_stats.restype = None
    lib.amr_hip_stats.argtypes = [H, C.POINTER(amr_gpu_stats)]
    lib.amr_hip_destroy.restype = None
    lib.amr_hip_destroy.argtypes = [H]
    lib.amr_hip_rhs.restype = C.c_int
    lib.amr_hip_rhs.argtypes = [H, c_double_p, C.c_uint64, c_double_p, C.c_char_p, C.c_size_t]
    lib.amr_hip_device_count.restype = C.c_int
    lib.amr_hip_device_count.argtypes = []
    lib.amr_hip_run.restype = C.c_int
    lib.amr_hip_run.argtypes = [H, C.c_double, C.c_int32, C.POINTER(amr_state_meta), C.c_char_p, C.c_size_t]
    lib.amr_hip_profile.restype = C.c_int
    lib.amr_hip_profile.argtypes = [H, C.c_double, C.c_int32, C.POINTER(amr_kernel_time), C.c_int32,
                                    C.POINTER(C.c_int32), C.c_char_p, C.c_size_t]
    lib.amr_hip_run_noalloc.restype = C.c_int
    lib.amr_hip_run_noalloc.argtypes = [C.POINTER(amr_mesh_view), C.POINTER(amr_params), c_double_p, C.c_uint64,
                                        C.POINTER(amr_state_meta), C.c_double, C.c_int64, C.c_double, C.c_int32,
                                        C.c_char_p, C.c_size_t]
    lib.amr_hip_copy_variants.restype = C.c_int
    lib.amr_hip_copy_variants.argtypes = [C.c_uint64, C.c_int32, c_double_p, C.c_int32, C.c_char_p, C.c_size_t]
    lib.amr_hip_copy_bandwidth.restype = C.c_int
    lib.amr_hip_copy_bandwidth.argtypes = [C.c_uint64, C.c_int32, c_double_p, c_double_p, C.c_char_p, C.c_size_t]
    lib.amr_hip_state.restype = C.c_int
    lib.amr_hip_state.argtypes = [H, C.POINTER(amr_state_meta)]
    lib.amr_hip_monitor.restype = C.c_int
    lib.amr_hip_monitor.argtypes = [H, c_double_p, c_double_p, C.c_double, C.c_double, c_double_p, C.c_char_p,
                                    C.c_size_t]
    lib.amr_hip_solve.restype = C.c_int
    lib.amr_hip_solve.argtypes = [H, C.POINTER(amr_steady_options), C.POINTER(amr_steady_result),
                                  C.POINTER(amr_steady_check), C.c_int32, C.c_char_p, C.c_size_t]
    _bind_batch_api(lib)
    _bind_ugrid_api(lib)
    lib.amr_hip_check_layout.restype = C.c_int
    lib.amr_hip_check_layout.argtypes = [C.POINTER(amr_mesh_view), C.POINTER(amr_params), C.c_int32, C.c_int32,
                                         C.c_char_p, C.c_size_t]
    if path is None:
        _LIB = lib
    return lib


def dptr(array: np.ndarray):
    """double* of a contiguous float64 array."""
    assert array.dtype == np.float64 and array.flags["C_CONTIGUOUS"]
    return array.ctypes.data_as(c_double_p)
