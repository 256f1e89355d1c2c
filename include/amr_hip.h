/* amr_hip.h -- C ABI of the MI355X (gfx950) cut-cell AMR time step library, libamr_hip.so.
 *
 * Three groups of entry points, all extern "C", plain pointers and sizes:
 *   amr_mesh_*   host mesher: static 2:1 quadtree cut-cell mesh, viscous and SA tables.
 *                Replaces static_amr2d_build / static_amr2d_prepare_viscous /
 *                static_amr2d_initialize (reference src/static_amr2d.h:72-76,
 *                src/static_amr2d_viscous.h:181).
 *   amr_hip_*    one device-resident solve; the five calls of the reference backend API
 *                src/static_amr2d_gpu.h:17-24 (create/advance/download/stats/destroy) plus a
 *                single residual evaluation that mirrors static_amr2d_rhs
 *                (src/static_amr2d.h:80-82).
 *   amr_batch_*  many meshes of uneven size advanced in lockstep in one device arena (no
 *                counterpart in the reference).
 *
 * Return codes follow the reference API: 0 ok, 1 error (message copied into err when err is
 * not NULL), 2 wall-clock budget expired with a resumable accepted state.
 * The element records below are byte-identical to the reference's C++ structs
 * (AmrCell2D 88 B, AmrFace2D 56 B, WallFragment2D 80 B, ...).
 */
#ifndef HIPAMR_AMR_HIP_H
#define HIPAMR_AMR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- element records (reference src/static_amr2d.h:37-45, src/euler2d.h:87-91,
 *      src/static_amr2d_viscous.h:128-130, src/static_amr2d_boundary.h:14-21) ---- */
typedef struct amr_cell {
  int32_t ix, iy, span, level;
  double x0, x1, y0, y1, volume, cx, cy;
  uint64_t wall_begin, wall_count;
} amr_cell;
typedef struct amr_face {
  int32_t left, right, axis, boundary; /* -1 = outside; normal along +axis; side 0..3 or -1 */
  double coordinate, lo, hi, length, centroid;
} amr_face;
typedef struct amr_wall {
  uint64_t owner_cell;
  int32_t source_edge_id, body_id, patch_id;
  double x0, y0, x1, y1, nx, ny, length; /* normal fluid -> solid */
} amr_wall;
typedef struct amr_grad_term { int32_t source, side; double x, y; } amr_grad_term;
typedef struct amr_wall_sample { int32_t donor; double distance; } amr_wall_sample;
typedef struct amr_face_geometry { double dx, dy, ex, ey, distance; } amr_face_geometry;
typedef struct amr_boundary {
  int32_t kind;
  double reference[4];
  double total_pressure, total_temperature, gas_constant;
  double direction[2];
  double outlet_pressure;
  double wall_velocity[2];
} amr_boundary;
typedef struct amr_transport {
  int32_t model, law, stabilization;
  double mu_ref, gas_constant, temperature_ref, sutherland, viscous_cfl;
  int32_t sa_model;
  double sa_nu_inf, sa_cfl;
  double wall_probe_min_h;
} amr_transport;

/* Time-step parameters: the fields of Params2D (src/euler2d.h:38-52) the path reads. */
typedef struct amr_params {
  double gamma, cfl;
  int32_t flux;    /* 0 HLLC, 1 AUSM+-up */
  int32_t limiter; /* must be 3 (first order) */
  double ausm_mco;
  int32_t bc[4];   /* legacy side kinds used where a boundary kind is -1 */
  double dt_fixed; /* > 0 fixed dt, 0 CFL */
  int32_t time_integrator; /* must be 0 (SSP-RK3) */
} amr_params;

/* Read-only view of one mesh: pointers into host memory owned by the caller. */
typedef struct amr_mesh_view {
  uint64_t n_cells, n_faces, n_walls;
  const amr_cell* cells;
  const amr_face* faces;
  const amr_wall* walls;
  const uint64_t* face_offsets;   /* n_cells + 1, cell -> face CSR */
  const uint64_t* cell_faces;
  const uint64_t* group_offsets;  /* n_cells + 1, SRD groups (group g seeded by cell g) */
  const uint64_t* group_members;
  const uint64_t* member_offsets; /* n_cells + 1, groups containing each cell */
  const uint64_t* member_groups;
  const double* group_volume;     /* n_cells */
  uint64_t small_cells;           /* > 0 enables SRD */
  int32_t agglomerate_small;
  amr_boundary boundaries[4];
  amr_transport transport;
  const uint64_t* visc_offsets;            /* n_cells + 1 (laminar) or NULL */
  const amr_grad_term* visc_terms;
  const amr_wall_sample* wall_samples;     /* n_walls */
  const amr_wall_sample* face_samples;     /* n_faces */
  const amr_face_geometry* face_geometry;  /* n_faces */
  const double* wall_distance;             /* n_cells (SA) or NULL */
} amr_mesh_view;

/* ---------------------------------------------------------------- host mesher ---- */
typedef struct amr_mesh amr_mesh; /* opaque: one built mesh in host memory */
typedef struct amr_box { double x0, x1, y0, y1; int32_t level; } amr_box;
typedef struct amr_build_options {
  int32_t nx, ny, levels, buffer;
  double xmin, xmax, ymin, ymax;
  uint64_t max_cells;
  int32_t agglomerate_small;
  int32_t n_boxes;
  const amr_box* boxes;
} amr_build_options;
typedef struct amr_mesh_info {
  uint64_t n_cells, n_faces, n_walls, n_cell_faces, n_group_members, n_visc_terms;
  uint64_t background_leaves, split_leaves, coarse_fine_faces, small_cells;
  uint64_t n_partition, n_leaf_components, n_agglomeration;
} amr_mesh_info;

/* Build the mesh (src/static_amr2d.cpp:229-358).  patches may be NULL. */
int amr_mesh_build(const amr_build_options* options, const double* px, const double* py,
                   const int32_t* patches, int32_t n_points, amr_mesh** out, char* err,
                   size_t err_len);
/* Set the four side boundaries and the transport model before prepare_viscous. */
int amr_mesh_set_physics(amr_mesh* mesh, const amr_boundary* boundaries4,
                         const amr_transport* transport);
/* Viscous/SA tables (src/static_amr2d_viscous.cpp:39-76, src/static_amr2d_sa.cpp:5-15). */
int amr_mesh_prepare_viscous(amr_mesh* mesh, const amr_params* params, char* err, size_t err_len);
int amr_mesh_get_view(const amr_mesh* mesh, amr_mesh_view* out);
int amr_mesh_get_info(const amr_mesh* mesh, amr_mesh_info* out);
/* Uniform state from a primitive (rho, u, v, p) (src/static_amr2d.cpp:360-367). */
int amr_mesh_initialize(const amr_mesh* mesh, const amr_params* params, const double* primitive4,
                        double* u, uint64_t u_len);
/* Audit tables: partition rows (ix, iy, span, level, first cell) as int32[5],
 * leaf_components rows as uint64[2], agglomeration rows as double[9]. */
int amr_mesh_copy_partition(const amr_mesh* mesh, int32_t* out, uint64_t rows);
int amr_mesh_copy_leaf_components(const amr_mesh* mesh, uint64_t* out, uint64_t rows);
int amr_mesh_copy_agglomeration(const amr_mesh* mesh, double* out, uint64_t rows);
void amr_mesh_destroy(amr_mesh* mesh);

/* ---------------------------------------------------- one device-resident solve ---- */
typedef struct amr_session amr_session; /* opaque (the reference's StaticAmrGpu) */
typedef struct amr_session_options {
  int32_t batch_trials; /* graph replays per host poll, 1..1024 */
  int32_t graph;        /* 1: replay a captured hipGraph of one step */
  int32_t event_timing; /* 1: hipEvent time per poll batch */
  int32_t precision;    /* 64 (FP64 state and fluxes) or 32 (FP32 state, FP64 control) */
} amr_session_options;
typedef struct amr_state_meta { double time; int64_t steps, rejected; } amr_state_meta;
/* Mirrors AmrGpuStats (src/static_amr2d_gpu.h:9-16). */
typedef struct amr_gpu_stats {
  uint64_t allocated_bytes, pinned_bytes, geometry_h2d_bytes, state_h2d_bytes;
  uint64_t control_h2d_bytes;
  uint64_t control_d2h_bytes, state_d2h_bytes, status_polls, full_downloads;
  double setup_seconds, advance_seconds, download_seconds;
  double advance_event_seconds;
  uint64_t event_batches, graph_nodes_per_trial;
} amr_gpu_stats;

/* static_amr_gpu_create (src/static_amr2d_gpu.h:17-18, src/static_amr2d_cuda.cu:274-325). */
int amr_hip_create(const amr_mesh_view* mesh, const amr_params* params, const double* u,
                   uint64_t u_len, const amr_state_meta* initial,
                   const amr_session_options* options, amr_session** out, char* err,
                   size_t err_len);
/* static_amr_gpu_advance (src/static_amr2d_gpu.h:19-21, src/static_amr2d_cuda.cu:327-353). */
int amr_hip_advance(amr_session* s, double tend, int64_t maxsteps, double wall_budget_seconds,
                    amr_state_meta* state, char* err, size_t err_len);
/* static_amr_gpu_download (src/static_amr2d_gpu.h:22, src/static_amr2d_cuda.cu:355-363). */
int amr_hip_download(amr_session* s, double* u, uint64_t u_len, amr_state_meta* state, char* err,
                     size_t err_len);
/* static_amr_gpu_stats / static_amr_gpu_destroy (src/static_amr2d_gpu.h:23-24). */
void amr_hip_stats(const amr_session* s, amr_gpu_stats* out);
void amr_hip_destroy(amr_session* s);
/* One residual evaluation at a host state, as static_amr2d_rhs (src/static_amr2d.cpp:369-377):
 * r gets nv * n_cells values.  Uses the session's mesh; leaves the session state untouched. */
int amr_hip_rhs(amr_session* s, const double* u, uint64_t u_len, double* r, char* err,
                size_t err_len);
/* Number of visible HIP devices (0 when no device or no driver). */
int amr_hip_device_count(void);
/* Host-only check of the device layout of n meshes (views[n], params[n]) at a launch tile of
 * 64, 128 or 256: every index a kernel dereferences stays inside its array and its mesh.
 * 0 consistent, 1 error.  No device is touched. */
int amr_hip_check_layout(const amr_mesh_view* views, const amr_params* params, int32_t n, int32_t tile, char* err,
                         size_t err_len);

#ifdef __cplusplus
}
#endif
#endif /* HIPAMR_AMR_HIP_H */
