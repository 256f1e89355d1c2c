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
 *   amr_batch_*  many meshes of uneven size advanced in lockstep in one device arena, slots
 *                admitted and released while others run (no counterpart in the reference).
 *   amr_driver_* the steady check loop of src/quick_airfoil.cpp:66-95 run over every slot of
 *                a batch on a host thread (begin / poll / running), for continuous batching.
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
  int32_t precision;    /* 64 (all FP64) or 32 (FP32 fluxes and work arrays; FP64 state, control, reductions) */
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
/* Time and counters of the last accepted state the host has seen (no device access). */
int amr_hip_state(const amr_session* s, amr_state_meta* state);
/* One residual evaluation at a host state, as static_amr2d_rhs (src/static_amr2d.cpp:369-377):
 * r gets nv * n_cells values.  Uses the session's mesh; leaves the session state untouched. */
int amr_hip_rhs(amr_session* s, const double* u, uint64_t u_len, double* r, char* err,
                size_t err_len);
/* Steady monitor at the accepted state (src/quick_airfoil.cpp:75-85), reduced in FP64 on the
 * device: out[0..4] raw residual norms R_v = sum_region V |L_v(u)| / sum_region V (v < nv; the
 * SA norm divided by nutilde_inf; 0 beyond nv), out[5..7] CL, CD, CM_z (q_inf = M^2/2, lift and
 * drag axes rotated by the angle of attack), out[8..10] fx, fy, mz (pressure minus viscous
 * wall shear, moment about `moment`).  region = {xmin, xmax, ymin, ymax} on cell centroids. */
int amr_hip_monitor(amr_session* s, const double* region4, const double* moment2, double mach, double aoa_degrees,
                    double* out11, char* err, size_t err_len);

/* Steady / time-accurate driver over one session (the loop of src/quick_airfoil.cpp:66-95):
 * advance to 0.01, then every `period`; after each target the monitor above; converged when
 * t >= min_time, max raw residual <= raw_tol and the relative changes of CL, CD, CM_z stay
 * <= load_tol for `hold` consecutive checks. */
typedef struct amr_steady_options {
  double mach, aoa_degrees;
  double moment[2];
  double region[4];
  double raw_tol, load_tol, min_time, period;
  int32_t hold;
  int32_t mode; /* 0 steady, 1 time (run to stop_time), 2 smoke (one check at 0.01) */
  double stop_time, budget_seconds;
} amr_steady_options;
typedef struct amr_steady_check {
  double time;
  int64_t steps, rejected;
  double elapsed;
  double residual[5];
  double cl, cd, cm, dcl, dcd, dcm;
  int32_t hold, pad;
} amr_steady_check;
enum {
  AMR_STOP_CONVERGED = 0, AMR_STOP_BUDGET = 1, AMR_STOP_TIME_CAP = 2, AMR_STOP_FAILURE = 3, AMR_STOP_SHORT = 4
};
typedef struct amr_steady_result {
  int32_t converged, failed, reason, hold, checks, pad;
  double time;
  int64_t steps, rejected;
  double residual[5];
  double cl, cd, cm;
  double elapsed_seconds;
} amr_steady_result;
/* history (may be NULL) receives up to history_capacity checks; result->checks counts all. */
int amr_hip_solve(amr_session* s, const amr_steady_options* options, amr_steady_result* result,
                  amr_steady_check* history, int32_t history_capacity, char* err, size_t err_len);

/* Run exactly `trials` step trials toward tend (graph replays, no intermediate poll), then
 * read the control record.  The timing primitive of the benchmark.  0 ok, 1 fatal. */
int amr_hip_run(amr_session* s, double tend, int32_t trials, amr_state_meta* state, char* err, size_t err_len);

/* Per-kernel device time: `trials` trials with direct launches and a hipEvent pair around
 * every kernel on the session stream, aggregated by kernel name (the solution advances). */
typedef struct amr_kernel_time { char name[32]; double total_ms; int64_t launches; } amr_kernel_time;
int amr_hip_profile(amr_session* s, double tend, int32_t trials, amr_kernel_time* out, int32_t capacity,
                    int32_t* count, char* err, size_t err_len);

/* One solve that allocates nothing.  No counterpart in the reference: for callers that cannot
 * allocate device memory (harnesses that forbid hipMalloc / hipFree, tight serving loops that
 * must not touch the allocator).  The same result as amr_hip_create (graph 0, the given
 * precision) + amr_hip_advance(tend, maxsteps, wall_budget_seconds) + amr_hip_download +
 * amr_hip_destroy, bit for bit: u (host, nv * n_cells) holds the initial state on entry and the
 * last accepted state on return, state carries time / steps / rejected in and out; returns 0,
 * 1 (error, or a fatal code with the last accepted state written back) or 2 (wall budget).
 * The arena lives in a static device array of the library, 64 MiB per device (about 70k cells
 * in FP64); a larger mesh is rejected with a message.  No device, pinned or graph allocation:
 * staging goes through pageable host memory on the null stream.  Every call on a device reuses
 * that device's array and overwrites it (nothing carries over from one call to the next); calls
 * that overlapped on one device would race for it, so calls from threads of one process take
 * turns behind a per-device lock. */
int amr_hip_run_noalloc(const amr_mesh_view* mesh, const amr_params* params, double* u, uint64_t u_len,
                        amr_state_meta* state, double tend, int64_t maxsteps, double wall_budget_seconds,
                        int32_t precision, char* err, size_t err_len);

/* STREAM-copy probe: `repeats` launches of a device copy of `bytes` (read + write) in each of
 * four kernel variants (16 B per lane; 4 x 16 B per lane; 4 x 16 B non-temporal; 8 x 16 B
 * non-temporal), gbps[v] = 2 * bytes / time (capacity >= 4).  The achievable-bandwidth reference
 * and the HBM counter calibration. */
int amr_hip_copy_variants(uint64_t bytes, int32_t repeats, double* gbps, int32_t capacity, char* err,
                          size_t err_len);
/* The best rate of amr_hip_copy_variants and its time per copy. */
int amr_hip_copy_bandwidth(uint64_t bytes, int32_t repeats, double* gbps, double* ms_per_copy, char* err,
                           size_t err_len);

/* ---- batch: many meshes of uneven size advanced together in one device arena -----------
 * No counterpart in the reference, which advances one mesh per process
 * (src/quick_airfoil.cpp:46-95).  The arena is a fixed set of slots; each slot owns
 * tile-aligned ranges of the cell, face and wall index spaces and of the ragged tables, sized
 * by its capacity.  Every slot has its own time, target, step size, counters and fatal code,
 * so a slot's trajectory does not depend on which other meshes share the arena (same
 * precision and tile give bitwise the same result as running it alone).  A trial of every
 * slot runs as up to three independent branches (slot m in branch m % 3), each a hipGraph
 * replayed on its own stream; the streams join at every poll.  Finished, failed and empty
 * slots are no-ops inside them. */
typedef struct amr_batch amr_batch;
typedef struct amr_slot_capacity {
  int32_t cells;   /* tile multiple, at least n_cells + 1 */
  int32_t faces, walls; /* tile multiples */
  int32_t csr, terms, groups, members; /* ragged table entries */
  int32_t pad;
} amr_slot_capacity;
typedef struct amr_batch_features { int32_t laminar, sa, small, visc_group; } amr_batch_features;
typedef struct amr_batch_options {
  int32_t precision;    /* 64 or 32 */
  int32_t tile;         /* launch tile: 64, 128 or 256 threads per block */
  int32_t graph;        /* 1: replay captured hipGraphs (one per branch) of one trial of every slot */
  int32_t batch_trials; /* trials per host poll, 1..1024 */
} amr_batch_options;
typedef struct amr_slot_state {
  double time, target;
  int64_t steps, rejected;
  int32_t n_cells;  /* 0: empty slot */
  int32_t finished; /* time >= target */
  int32_t fatal;    /* 0, or the fatal code of the reference (2 nonphysical, 3 maxsteps, 4 dt underflow,
                       5 rejections); the slot keeps its last accepted state */
  int32_t pad;
} amr_slot_state;

/* Smallest slot that holds the mesh at this tile, and the kernels it needs. */
int amr_batch_capacity(const amr_mesh_view* mesh, int32_t tile, amr_slot_capacity* capacity,
                       amr_batch_features* features);
/* Device bytes of an arena of n slots (its single allocation), for HBM budget admission. */
uint64_t amr_batch_bytes(const amr_slot_capacity* capacities, int32_t n, const amr_batch_features* features,
                         int32_t precision, int32_t tile);
/* One slot per mesh, each sized to its mesh and admitted with state u[m] (nv * n_cells doubles)
 * and initial[m] (NULL: t = 0, counters 0). */
int amr_batch_create(const amr_mesh_view* meshes, const amr_params* params, const double* const* u,
                     const amr_state_meta* initial, int32_t n, const amr_batch_options* options, amr_batch** out,
                     char* err, size_t err_len);
/* n empty slots of the given capacities; the kernel set is fixed by `features` (a slot accepts
 * meshes that need no more: same SA-neg flag, laminar/small/grouped only if enabled). */
int amr_batch_create_slots(const amr_slot_capacity* capacities, int32_t n, const amr_batch_features* features,
                           const amr_batch_options* options, amr_batch** out, char* err, size_t err_len);
/* Place a mesh into an empty or occupied slot (the previous mesh is dropped); the slot starts
 * finished at initial->time until its next target. */
int amr_batch_admit(amr_batch* b, int32_t slot, const amr_mesh_view* mesh, const amr_params* params,
                    const double* u, uint64_t u_len, const amr_state_meta* initial, char* err, size_t err_len);
/* Empty a slot. */
int amr_batch_release(amr_batch* b, int32_t slot, char* err, size_t err_len);
/* Target time and step cap of every slot (n entries each; maxsteps NULL: no cap).  Empty and
 * failed slots ignore theirs. */
int amr_batch_set_targets(amr_batch* b, const double* targets, const int64_t* maxsteps, char* err,
                          size_t err_len);
/* Replay trials until every slot is finished or failed (until = 0), until at least one slot
 * that was running at the call has finished or failed (until = 1), or, grouped (until = 2),
 * until an eighth of those slots have ended or min(8, running / 32) polls have passed since
 * the first ended, polling the control records every batch_trials trials.  Ended slots wait
 * idle, so grouping trades a short wait for fewer checks.  Polls are pipelined: the next
 * batch_trials trials are already queued when a poll is read, so slots still running when the
 * condition is met advance up to batch_trials more trials (never past their targets or step
 * caps) before the call returns.  states (may be NULL) receives every slot as of the return.
 * 0 done, 2 wall budget expired (0 = no budget), 1 device error. */
int amr_batch_advance(amr_batch* b, int32_t until, double wall_budget_seconds, amr_slot_state* states, char* err,
                      size_t err_len);
/* Exactly `trials` trials, then one poll. */
int amr_batch_run(amr_batch* b, int32_t trials, amr_slot_state* states, char* err, size_t err_len);
/* Launch shape of the trials (no counterpart in the reference; a batch tuning knob): the flux
 * launches run flux_spread threads per cell (1 or 2; tile * spread <= 256), and while at most
 * few_slots slots are running (0: never) few_flux_spread instead, so that a nearly drained batch,
 * whose trials are bound by latency, spreads each tile's face items over more threads.  Results
 * are bitwise independent of the shape.  Recaptures the trial graphs.  0 ok, 1 invalid or error. */
int amr_batch_set_launch_shape(amr_batch* b, int32_t flux_spread, int32_t few_slots, int32_t few_flux_spread, char* err,
                               size_t err_len);
/* Slot states of the last poll (no device access). */
int amr_batch_slots(const amr_batch* b, amr_slot_state* states);
int32_t amr_batch_size(const amr_batch* b);
/* Accepted state of one slot (nv * n_cells doubles). */
int amr_batch_download(amr_batch* b, int32_t slot, double* u, uint64_t u_len, amr_state_meta* state, char* err,
                       size_t err_len);
/* Steady monitor of amr_hip_monitor for every slot with mask[m] != 0, in one residual pass:
 * frames[6m..6m+5] = region xmin, xmax, ymin, ymax and moment point; out[11m..11m+10] as
 * amr_hip_monitor; status[m] = 0 ok, 1 nonphysical state at the monitor, 2 empty region,
 * 3 not evaluated (mask 0, empty or failed slot). */
int amr_batch_monitor(amr_batch* b, const int32_t* mask, const double* frames, const double* mach,
                      const double* aoa_degrees, double* out, int32_t* status, char* err, size_t err_len);
/* Per-kernel device time over `trials` directly launched trials of the whole arena on one
 * stream, unbranched (as amr_hip_profile). */
int amr_batch_profile(amr_batch* b, int32_t trials, amr_kernel_time* out, int32_t capacity, int32_t* count,
                      char* err, size_t err_len);
void amr_batch_stats(const amr_batch* b, amr_gpu_stats* out);
void amr_batch_destroy(amr_batch* b);

/* Steady driver over a batch: every slot runs its own check loop (the loop of amr_hip_solve,
 * src/quick_airfoil.cpp:66-95) with its own options; slots finish independently. */
typedef struct amr_batch_driver amr_batch_driver;
int amr_driver_create(amr_batch* b, amr_batch_driver** out, char* err, size_t err_len);
/* Start a steady run on an admitted slot from its current state (its wall budget starts now). */
int amr_driver_begin(amr_batch_driver* d, int32_t slot, const amr_steady_options* options, char* err,
                     size_t err_len);
/* Advance the running slots until at least one run ends or wall_budget_seconds (0: none)
 * passes.  Ended runs: their slot ids in done[0..*n_done-1], results[slot] filled.  0 ok,
 * 2 budget expired with no run ended, 1 device error. */
int amr_driver_poll(amr_batch_driver* d, double wall_budget_seconds, amr_steady_result* results, int32_t* done,
                    int32_t* n_done, char* err, size_t err_len);
/* Number of runs in progress. */
int32_t amr_driver_running(const amr_batch_driver* d);
void amr_driver_destroy(amr_batch_driver* d);
/* Convenience: a steady run on every occupied slot (options[m] per slot), polled to the end. */
int amr_batch_solve(amr_batch* b, const amr_steady_options* options, amr_steady_result* results, char* err,
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
