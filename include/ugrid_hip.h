/* ugrid_hip.h -- C ABI of the batched tensor-product-grid Euler solver on gfx950: the
 * reference's legacy 2-D solver family (src/euler2d.h:38-225), whose GPU backend is
 * solver2d_run_cuda (src/euler2d.h:211, src/euler2d_cuda.cu:530-708).
 *
 *   ugrid_axis      grid2d_stretch_faces + the axis geometry of solver2d_init
 *                   (src/euler2d.cpp:38-105): face coordinates, cell centres and widths with
 *                   two ghost layers per side.
 *   ugrid_create    upload one batch (ncase cases sharing the grid; each its own state, time,
 *                   last step and AUSM+-up reference Mach number) -- the set-up part of
 *                   solver2d_run_cuda (src/euler2d_cuda.cu:558-597).
 *   ugrid_run       advance every case to tend (src/euler2d_cuda.cu:599-693): CFL or fixed dt
 *                   per case, SSP-RK3; a candidate with a nonphysical cell in any case halves
 *                   every case's step and is retried (at most 12 times).  Returns 0 (every case
 *                   at tend), 1 (maxsteps steps taken in this call), 2 (nonphysical candidate
 *                   after 12 retries; err holds the message), 3 (error).  With time_integrator
 *                   1, backward Euler by pseudo-time defect relaxation (src/euler2d_cuda.cu:
 *                   370-447): 2 also for an invalid initial state or a non-finite defect, 3 for
 *                   invalid controls or a physical dt that cannot advance, 4 when a step does
 *                   not converge within implicit_maxiter updates.
 *   ugrid_rhs       the residual at a state, boundary ghosts filled from its interior first
 *                   (solver2d_apply_bc + solver2d_residual, src/euler2d.cpp:249-489, as the GPU
 *                   kernels of src/euler2d_cuda.cu:153-258 compute it).
 *   ugrid_download  state, times and last steps (solver2d_run_cuda's write-back, :694-707).
 *   ugrid_destroy.
 *
 * Method as the reference's GPU backend: first order (limiter 3 only), HLLC or AUSM+-up,
 * boundary kinds 0 periodic, 1 transmissive, 2 reflect per side (xmin, xmax, ymin, ymax), FP64.
 * State layout as the reference (src/euler2d.h:137-140):
 * U[((case * 4 + var) * my + j + 2) * mx + i + 2], var = rho, rho u, rho v, rho E,
 * mx = nx + 4, my = ny + 4.  Ghost values on entry are not read; on download they are what the
 * reference's device array holds after the same call (the boundary fill of the state at the
 * start of the last step begun, or the entry values when no step began).
 */
#ifndef HIPAMR_UGRID_HIP_H
#define HIPAMR_UGRID_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "amr_hip.h" /* amr_wall: the WallFragment2D record */

#ifdef __cplusplus
extern "C" {
#endif

/* Stretching of one axis (src/euler2d.h:27-32): 0 uniform, 1 geometric (adjacent-width ratio
 * param), 2 tanh at both ends, 3 sinh at the centre (param delta).  xf gets n + 1 face
 * coordinates, xc and dx n + 4 centres and widths (two ghosts per side; periodic ghosts are
 * translated, others mirrored).  Returns 0, or 1 with a message. */
int ugrid_axis(int32_t n, double a, double b, int32_t stretch_type, double stretch_param, int32_t bc_lo,
               int32_t bc_hi, double* xf, double* xc, double* dx, char* err, size_t err_len);

/* The solver fields of Params2D (src/euler2d.h:38-52). */
typedef struct ugrid_params {
  double gamma, cfl;
  int32_t flux;             /* 0 HLLC, 1 AUSM+-up */
  int32_t limiter;          /* must be 3 (first order) */
  double ausm_mco;          /* default reference Mach number of the cases */
  int32_t bc[4];            /* xmin, xmax, ymin, ymax: 0 periodic, 1 transmissive, 2 reflect */
  double dt_fixed;          /* > 0 fixed physical dt, else CFL */
  int32_t time_integrator;  /* 0 SSP-RK3, 1 backward Euler */
  int32_t implicit_maxiter;
  double pseudo_cfl, implicit_rtol, implicit_atol;
} ugrid_params;

/* Cut-cell geometry of a grid (the arrays of Cut2D, src/euler2d.h:98-117, same layouts and
 * meanings; mx * my cell arrays include the ghosts).  The device backend reads nsmall,
 * nbad_nbhd, vol, flag, axl, ayl, wax, way, nbrcount, smallpos, scnt, smem and svhat. */
typedef struct ugrid_cut_view {
  int32_t nsmall, nbad_nbhd, nsolid, ncut;
  uint64_t n_walls;
  const double *vol, *cx, *cy, *flag;   /* mx * my: fluid area (0 solid), centroid, 0 full / 1 cut / 2 solid */
  const double *axl, *axy;              /* (nx + 1) * ny x-face apertures and their y centroids, j * (nx + 1) + i */
  const double *ayl, *ayx;              /* nx * (ny + 1) y-face apertures and their x centroids, j * nx + i */
  const double *wax, *way, *wcx, *wcy;  /* mx * my wall vector (into the solid, times length), wall centroid */
  const double *nbrcount, *smallpos;    /* mx * my SRD neighbourhood count, position in the small-cell list or -1 */
  const double *sidx, *scnt, *smem;     /* nsmall cell indices, member counts, 25 * nsmall members */
  const double *svhat, *sxhat, *syhat;  /* nsmall weighted volumes and centroids */
  const amr_wall* walls;                /* n_walls wall fragments (normal fluid -> solid) */
} ugrid_cut_view;

typedef struct ugrid_setup {
  int32_t nx, ny, ncase;
  int32_t pad;
  const double* dx;     /* nx + 4 cell widths, ghosts included (ugrid_axis) */
  const double* dy;     /* ny + 4 */
  ugrid_params params;
  const double* mco;    /* ncase AUSM+-up reference Mach numbers, or NULL for params.ausm_mco */
  const ugrid_cut_view* cut;  /* NULL: regular grid */
} ugrid_setup;

/* Cut-cell geometry and SRD neighbourhoods of a closed solid polygon on the grid
 * (cut2d_build_polygon, src/cutcell2d.cpp:247-503; bit-identical arrays).  axes: nx, ny and the
 * ugrid_axis outputs of both axes.  patches: one patch id per edge or NULL.  Returns 0, 1 (an
 * invalid polygon), 2 (the body within three cells of the domain boundary) or 3 (an edge on a
 * grid line), as the reference does. */
typedef struct ugrid_axes {
  int32_t nx, ny;
  const double *xf, *xc, *dx;  /* nx + 1, nx + 4, nx + 4 */
  const double *yf, *yc, *dy;  /* ny + 1, ny + 4, ny + 4 */
} ugrid_axes;
typedef struct ugrid_cut ugrid_cut;
int ugrid_cut_build(const ugrid_axes* axes, const double* px, const double* py, int32_t n, const int32_t* patches,
                    ugrid_cut** out, char* err, size_t err_len);
int ugrid_cut_get(const ugrid_cut* cut, ugrid_cut_view* view);
void ugrid_cut_destroy(ugrid_cut* cut);

typedef struct ugrid_session ugrid_session;

typedef struct ugrid_meta {
  int64_t nsteps, nreject; /* accepted steps and rejected candidates since create */
} ugrid_meta;

/* u: ncase * 4 * mx * my doubles; t and dt: ncase times and last steps (NULL: all zero). */
int ugrid_create(const ugrid_setup* setup, const double* u, uint64_t u_len, const double* t, const double* dt,
                 ugrid_session** out, char* err, size_t err_len);
int ugrid_run(ugrid_session* s, double tend, int64_t maxsteps, ugrid_meta* meta, char* err, size_t err_len);
/* u gets the state (ghosts included), t the ncase times and dt the ncase last accepted steps;
 * t and dt may be NULL. */
int ugrid_download(ugrid_session* s, double* u, uint64_t u_len, double* t, double* dt, char* err, size_t err_len);
/* r gets the residual of u (same layout; ghost entries zero). */
int ugrid_rhs(ugrid_session* s, const double* u, uint64_t u_len, double* r, char* err, size_t err_len);
void ugrid_destroy(ugrid_session* s);

/* 1 when a HIP device is usable (reason: its name and architecture), else 0 (reason: why);
 * the check of solver2d_cuda_available (src/euler2d_cuda.cu:510-528). */
int ugrid_device_available(char* reason, size_t reason_len);

/* Backward Euler (time_integrator 1; src/euler2d_cuda.cu:370-447): pseudo-time updates and
 * candidate retries since create, convergence checks, the latest check's largest first norm,
 * norm and tolerance over the cases (ImplicitStats2D, src/euler2d.h:118-121), and why the last
 * run stopped (0 none, 1 invalid initial state, 2 candidate retries, 3 non-finite defect,
 * 4 iteration limit, 5 physical dt). */
typedef struct ugrid_implicit_info {
  int64_t updates, retries, checks, reason;
  double initial_defect, final_defect, tolerance;
} ugrid_implicit_info;
int ugrid_implicit_stats(const ugrid_session* s, ugrid_implicit_info* out);

/* Sizes and field offsets of the Solver2D view behind cfd::solver2d_run_cuda (euler2d_gpu.h):
 * Solver2D, Grid2D, Array, Params2D, ImplicitStats2D, then grid, grid.{nx, ng, mx, my, xmin,
 * xf, yf, xc, dx, yc, dy}, par, ncase, U, U1, U2, R, W, SX, SY, FX, FY, t, dt, mco, nsteps,
 * nreject, implicit, ib, cut; then the size of Cut2D and the offsets of nsmall, nbad_nbhd, vol,
 * flag, axl, ayl, wax, way, nbrcount, smallpos, scnt, smem, svhat.  Writes min(capacity, 50)
 * values, returns 50. */
int ugrid_solver2d_layout(int64_t* out, int32_t capacity);

/* Average duration in microseconds of each RK stage kernel (stage_us[0..2]) over `trials`
 * trials launched directly with HIP events around every kernel; the batch is advanced by
 * those trials (towards t = infinity, no step cap).  For the bench. */
int ugrid_profile(ugrid_session* s, int32_t trials, double* stage_us, char* err, size_t err_len);

/* Rows per wave strip and the launch shape of the stage kernels (for tests and the bench). */
typedef struct ugrid_launch_info {
  int32_t rows, strips, blocks, threads;
} ugrid_launch_info;
int ugrid_get_launch(const ugrid_session* s, ugrid_launch_info* out);

#ifdef __cplusplus
}
#endif

#endif
