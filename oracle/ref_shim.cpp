// ref_shim.cpp -- extern "C" handles over the reference's own CPU code, used only by tests to
// pin this project's mesher and oracle against the real reference in the build container.
//
// Compiled (see oracle/Makefile) together with the reference sources where they lie; the
// reference is never copied into this repository.  Entry points mirror amr_mesh_* of
// include/amr_hip.h with a ref_ prefix, plus the reference's CPU time step:
//   ref_rhs          static_amr2d_rhs        (src/static_amr2d.cpp:369-377)
//   ref_run          static_amr2d_run        (src/static_amr2d.cpp:384-432)
//   ref_redistribute static_amr2d_redistribute (src/static_amr2d.cpp:379-382)
//   ref_wall_flux    static_amr2d_viscous_diagnostics (src/static_amr2d_viscous.cpp:137-142)
//   ref_boundary_state / ref_boundary_valid
//                    amr_bc_state / amr_bc_valid (src/static_amr2d_boundary.h:32-102)
// and, for the legacy tensor-product grid solver (include/ugrid_hip.h):
//   ref_grid_geometry  solver2d_init's grid (src/euler2d.cpp:38-151)
//   ref_grid_apply_bc  solver2d_apply_bc (src/euler2d.cpp:221-247)
//   ref_grid_rhs       solver2d_apply_bc + solver2d_residual (src/euler2d.cpp:249-489)
//   ref_grid_run       solver2d_run, the CPU SSP-RK3 driver (src/euler2d.cpp:595-698)
//   ref_flux           flux_hllc_t / flux_ausmpup_t (src/flux.h:41-176)
#include <cstddef>
#include <cstring>
#include <memory>
#include <string>

#include "amr_hip.h"
#include "ugrid_hip.h"
#include "euler2d.h"
#include "flux.h"
#include "static_amr2d.h"
#include "static_amr2d_boundary.h"
#include "static_amr2d_sa.h"

using namespace cfd;

struct ref_mesh {
  StaticAmrMesh2D mesh;
};

static int ref_report(char* err, size_t len, const std::string& message) {
  if (err && len) {
    const size_t k = message.size() < len - 1 ? message.size() : len - 1;
    std::memcpy(err, message.data(), k);
    err[k] = 0;
  }
  return 1;
}

static Params2D ref_params(const amr_params& c) {
  Params2D p = params2d_default();
  p.gamma = c.gamma;
  p.cfl = c.cfl;
  p.flux = c.flux;
  p.limiter = c.limiter;
  p.ausm_mco = c.ausm_mco;
  for (int k = 0; k < 4; ++k) p.bc[k] = c.bc[k];
  p.dt_fixed = c.dt_fixed;
  p.time_integrator = c.time_integrator;
  return p;
}

extern "C" {

int ref_mesh_build(const amr_build_options* options, const double* px, const double* py, const int32_t* patches,
                   int32_t n_points, ref_mesh** out, char* err, size_t err_len) {
  if (!out || !options) return 1;
  *out = nullptr;
  StaticAmrOptions2D o;
  o.nx = options->nx;
  o.ny = options->ny;
  o.levels = options->levels;
  o.buffer = options->buffer;
  o.xmin = options->xmin;
  o.xmax = options->xmax;
  o.ymin = options->ymin;
  o.ymax = options->ymax;
  o.max_cells = options->max_cells;
  o.agglomerate_small = options->agglomerate_small != 0;
  for (int32_t b = 0; b < options->n_boxes; ++b)
    o.boxes.push_back({options->boxes[b].x0, options->boxes[b].x1, options->boxes[b].y0, options->boxes[b].y1,
                       options->boxes[b].level});
  std::vector<double> x(px, px + n_points), y(py, py + n_points);
  std::vector<int> pt;
  if (patches) pt.assign(patches, patches + n_points);
  ref_mesh* m = new ref_mesh;
  std::string error;
  if (static_amr2d_build(o, x, y, pt, &m->mesh, &error)) {
    delete m;
    return ref_report(err, err_len, error);
  }
  *out = m;
  return 0;
}

int ref_mesh_set_physics(ref_mesh* m, const amr_boundary* boundaries4, const amr_transport* transport) {
  if (!m) return 1;
  if (boundaries4) std::memcpy(m->mesh.boundaries, boundaries4, 4 * sizeof(amr_boundary));
  if (transport) std::memcpy(&m->mesh.transport, transport, sizeof(amr_transport));
  return 0;
}

int ref_mesh_prepare_viscous(ref_mesh* m, const amr_params* params, char* err, size_t err_len) {
  if (!m || !params) return 1;
  std::string error;
  if (static_amr2d_prepare_viscous(&m->mesh, ref_params(*params), &error)) return ref_report(err, err_len, error);
  return 0;
}

int ref_mesh_get_view(const ref_mesh* mm, amr_mesh_view* v) {
  if (!mm || !v) return 1;
  const StaticAmrMesh2D& m = mm->mesh;
  std::memset(v, 0, sizeof(*v));
  v->n_cells = m.cells.size();
  v->n_faces = m.faces.size();
  v->n_walls = m.walls.size();
  v->cells = reinterpret_cast<const amr_cell*>(m.cells.data());
  v->faces = reinterpret_cast<const amr_face*>(m.faces.data());
  v->walls = reinterpret_cast<const amr_wall*>(m.walls.data());
  v->face_offsets = reinterpret_cast<const uint64_t*>(m.face_offsets.data());
  v->cell_faces = reinterpret_cast<const uint64_t*>(m.cell_faces.data());
  v->group_offsets = reinterpret_cast<const uint64_t*>(m.group_offsets.data());
  v->group_members = reinterpret_cast<const uint64_t*>(m.group_members.data());
  v->member_offsets = reinterpret_cast<const uint64_t*>(m.member_offsets.data());
  v->member_groups = reinterpret_cast<const uint64_t*>(m.member_groups.data());
  v->group_volume = m.group_volume.data();
  v->small_cells = m.small_cells;
  v->agglomerate_small = m.options.agglomerate_small ? 1 : 0;
  std::memcpy(v->boundaries, m.boundaries, sizeof(v->boundaries));
  std::memcpy(&v->transport, &m.transport, sizeof(v->transport));
  if (!m.viscous.offsets.empty()) {
    v->visc_offsets = reinterpret_cast<const uint64_t*>(m.viscous.offsets.data());
    v->visc_terms = reinterpret_cast<const amr_grad_term*>(m.viscous.terms.data());
  }
  if (!m.viscous.wall_samples.empty()) v->wall_samples = reinterpret_cast<const amr_wall_sample*>(m.viscous.wall_samples.data());
  if (!m.viscous.face_samples.empty()) v->face_samples = reinterpret_cast<const amr_wall_sample*>(m.viscous.face_samples.data());
  if (!m.viscous.face_geometry.empty())
    v->face_geometry = reinterpret_cast<const amr_face_geometry*>(m.viscous.face_geometry.data());
  if (!m.viscous.wall_distance.empty()) v->wall_distance = m.viscous.wall_distance.data();
  return 0;
}

int ref_mesh_get_info(const ref_mesh* mm, amr_mesh_info* out) {
  if (!mm || !out) return 1;
  const StaticAmrMesh2D& s = mm->mesh;
  out->n_cells = s.cells.size();
  out->n_faces = s.faces.size();
  out->n_walls = s.walls.size();
  out->n_cell_faces = s.cell_faces.size();
  out->n_group_members = s.group_members.size();
  out->n_visc_terms = s.viscous.terms.size();
  out->background_leaves = s.background_leaves;
  out->split_leaves = s.split_leaves;
  out->coarse_fine_faces = s.coarse_fine_faces;
  out->small_cells = s.small_cells;
  out->n_partition = s.partition.size();
  out->n_leaf_components = s.leaf_components.size();
  out->n_agglomeration = s.agglomeration.size();
  return 0;
}

int ref_mesh_initialize(const ref_mesh* m, const amr_params* params, const double* primitive4, double* u,
                        uint64_t u_len) {
  if (!m || !params) return 1;
  StaticAmrState2D s;
  if (static_amr2d_initialize(m->mesh, ref_params(*params), {primitive4[0], primitive4[1], primitive4[2], primitive4[3]}, &s))
    return 1;
  if (s.u.size() != u_len) return 1;
  std::memcpy(u, s.u.data(), u_len * sizeof(double));
  return 0;
}

int ref_mesh_copy_partition(const ref_mesh* m, int32_t* out, uint64_t rows) {
  if (!m || rows != m->mesh.partition.size()) return 1;
  for (size_t k = 0; k < rows; ++k)
    for (int j = 0; j < 5; ++j) out[5 * k + j] = m->mesh.partition[k][j];
  return 0;
}

int ref_mesh_copy_leaf_components(const ref_mesh* m, uint64_t* out, uint64_t rows) {
  if (!m || rows != m->mesh.leaf_components.size()) return 1;
  for (size_t k = 0; k < rows; ++k) {
    out[2 * k] = m->mesh.leaf_components[k][0];
    out[2 * k + 1] = m->mesh.leaf_components[k][1];
  }
  return 0;
}

int ref_mesh_copy_agglomeration(const ref_mesh* m, double* out, uint64_t rows) {
  if (!m || rows != m->mesh.agglomeration.size()) return 1;
  for (size_t k = 0; k < rows; ++k)
    for (int j = 0; j < 9; ++j) out[9 * k + j] = m->mesh.agglomeration[k][j];
  return 0;
}

void ref_mesh_destroy(ref_mesh* m) { delete m; }

int ref_rhs(const ref_mesh* m, const amr_params* params, const double* u, uint64_t u_len, double* r, uint64_t r_len) {
  if (!m || !params) return 1;
  std::vector<double> state(u, u + u_len), out;
  if (static_amr2d_rhs(m->mesh, ref_params(*params), state, &out)) return 1;
  if (out.size() != r_len) return 2;
  std::memcpy(r, out.data(), r_len * sizeof(double));
  return 0;
}

int ref_run(const ref_mesh* m, const amr_params* params, double tend, int64_t maxsteps, double* u, uint64_t u_len,
            amr_state_meta* meta, char* err, size_t err_len) {
  if (!m || !params || !meta) return 1;
  StaticAmrState2D s;
  s.u.assign(u, u + u_len);
  s.time = meta->time;
  s.steps = meta->steps;
  s.rejected = meta->rejected;
  std::string error;
  const int status = static_amr2d_run(m->mesh, ref_params(*params), tend, maxsteps, &s, &error);
  std::memcpy(u, s.u.data(), u_len * sizeof(double));
  meta->time = s.time;
  meta->steps = s.steps;
  meta->rejected = s.rejected;
  if (status) return ref_report(err, err_len, error);
  return 0;
}

int ref_redistribute(const ref_mesh* m, double* u, uint64_t u_len) {
  if (!m) return 1;
  std::vector<double> state(u, u + u_len);
  if (static_amr2d_redistribute(m->mesh, &state)) return 1;
  std::memcpy(u, state.data(), u_len * sizeof(double));
  return 0;
}

// Viscous wall flux (4 * n_walls, already times fragment length) at a conserved state.
int ref_wall_flux(const ref_mesh* m, const amr_params* params, const double* u, uint64_t u_len, double* wall_flux,
                  uint64_t len) {
  if (!m || !params) return 1;
  std::vector<double> state(u, u + u_len);
  AmrViscousFields2D fields;
  if (static_amr2d_viscous_diagnostics(m->mesh, ref_params(*params), state, &fields)) return 1;
  if (fields.wall_flux.size() != len) return 2;
  std::memcpy(wall_flux, fields.wall_flux.data(), len * sizeof(double));
  return 0;
}

// Exterior Riemann state of one side (1 = valid state, 0 = physical/configuration failure).
int ref_boundary_state(const amr_boundary* b, int legacy, double gamma, int side, const double* wi, double* wb) {
  static_assert(sizeof(amr_boundary) == sizeof(AmrBoundary2D), "boundary record layout");
  AmrBoundary2D rb;
  std::memcpy(&rb, b, sizeof(rb));
  return amr_bc_state(rb, legacy, gamma, side, wi, wb) ? 1 : 0;
}

int ref_boundary_valid(const amr_boundary* b, int legacy, int side) {
  AmrBoundary2D rb;
  std::memcpy(&rb, b, sizeof(rb));
  return amr_bc_valid(rb, legacy, side) ? 1 : 0;
}


// ------------------------------------------------------------ legacy grid solver --
static Params2D ref_grid_params(const ugrid_params& c, const int* stretch_type, const double* stretch_param) {
  Params2D p = params2d_default();
  p.gamma = c.gamma;
  p.cfl = c.cfl;
  p.flux = c.flux;
  p.limiter = c.limiter;
  p.ausm_mco = c.ausm_mco;
  for (int k = 0; k < 4; ++k) p.bc[k] = c.bc[k];
  p.dt_fixed = c.dt_fixed;
  p.time_integrator = c.time_integrator;
  p.pseudo_cfl = c.pseudo_cfl;
  p.implicit_maxiter = c.implicit_maxiter;
  p.implicit_rtol = c.implicit_rtol;
  p.implicit_atol = c.implicit_atol;
  p.sx.type = stretch_type ? stretch_type[0] : STRETCH_UNIFORM;
  p.sx.param = stretch_param ? stretch_param[0] : 0.0;
  p.sy.type = stretch_type ? stretch_type[1] : STRETCH_UNIFORM;
  p.sy.param = stretch_param ? stretch_param[1] : 0.0;
  return p;
}

// box = xmin, xmax, ymin, ymax
static std::unique_ptr<Solver2D> ref_grid_solver(int nx, int ny, const double* box, int ncase, const ugrid_params* p,
                                                 const int* stretch_type, const double* stretch_param) {
  std::unique_ptr<Solver2D> s(new Solver2D());
  solver2d_init(s.get(), nx, ny, box[0], box[1], box[2], box[3], ncase, ref_grid_params(*p, stretch_type, stretch_param));
  return s;
}

int ref_grid_geometry(int nx, int ny, const double* box, const ugrid_params* p, const int* stretch_type,
                      const double* stretch_param, double* xf, double* xc, double* dx, double* yf, double* yc,
                      double* dy) {
  auto s = ref_grid_solver(nx, ny, box, 1, p, stretch_type, stretch_param);
  const Grid2D& g = s->grid;
  std::memcpy(xf, g.xf.data(), g.xf.size() * sizeof(double));
  std::memcpy(xc, g.xc.data(), g.xc.size() * sizeof(double));
  std::memcpy(dx, g.dx.data(), g.dx.size() * sizeof(double));
  std::memcpy(yf, g.yf.data(), g.yf.size() * sizeof(double));
  std::memcpy(yc, g.yc.data(), g.yc.size() * sizeof(double));
  std::memcpy(dy, g.dy.data(), g.dy.size() * sizeof(double));
  return 0;
}

void ref_grid_apply_bc(int nx, int ny, int ncase, const ugrid_params* p, double* u) {
  const double box[4] = {0.0, 1.0, 0.0, 1.0};
  auto s = ref_grid_solver(nx, ny, box, ncase, p, nullptr, nullptr);
  std::memcpy(s->U.data(), u, s->U.size() * sizeof(double));
  solver2d_apply_bc(s.get(), s->U.data());
  std::memcpy(u, s->U.data(), s->U.size() * sizeof(double));
}

int ref_grid_rhs(int nx, int ny, const double* box, int ncase, const ugrid_params* p, const int* stretch_type,
                 const double* stretch_param, const double* mco, const double* u, double* r) {
  auto s = ref_grid_solver(nx, ny, box, ncase, p, stretch_type, stretch_param);
  if (mco)
    for (int c = 0; c < ncase; ++c) solver2d_set_mco(s.get(), c, mco[c]);
  std::memcpy(s->U.data(), u, s->U.size() * sizeof(double));
  solver2d_apply_bc(s.get(), s->U.data());
  solver2d_residual(s.get(), s->U.data(), s->R.data(), s->W.data(), s->SX.data(), s->SY.data(), s->FX.data(),
                    s->FY.data());
  std::memcpy(r, s->R.data(), s->R.size() * sizeof(double));
  return 0;
}

int ref_grid_run(int nx, int ny, const double* box, int ncase, const ugrid_params* p, const int* stretch_type,
                 const double* stretch_param, const double* mco, double* u, double* t, double* dt, double tend,
                 int64_t maxsteps, int64_t* nsteps, int64_t* nreject) {
  auto s = ref_grid_solver(nx, ny, box, ncase, p, stretch_type, stretch_param);
  for (int c = 0; c < ncase; ++c) {
    if (mco) solver2d_set_mco(s.get(), c, mco[c]);
    s->t[c] = t[c];
    s->dt[c] = dt[c];
  }
  std::memcpy(s->U.data(), u, s->U.size() * sizeof(double));
  const int status = solver2d_run(s.get(), tend, static_cast<long>(maxsteps));
  std::memcpy(u, s->U.data(), s->U.size() * sizeof(double));
  for (int c = 0; c < ncase; ++c) {
    t[c] = s->t[c];
    dt[c] = s->dt[c];
  }
  *nsteps = s->nsteps;
  *nreject = s->nreject;
  return status;
}

// solver2d_run_implicit (src/implicit2d.cpp:104-207), the CPU backward-Euler driver.  counts:
// updates, retries (in/out); defects: initial, final, tolerance (out).
int ref_grid_run_implicit(int nx, int ny, const double* box, int ncase, const ugrid_params* p, const int* stretch_type,
                          const double* stretch_param, const double* mco, double* u, double* t, double* dt,
                          double tend, int64_t maxsteps, int64_t* nsteps, int64_t* counts, double* defects) {
  auto s = ref_grid_solver(nx, ny, box, ncase, p, stretch_type, stretch_param);
  for (int c = 0; c < ncase; ++c) {
    if (mco) solver2d_set_mco(s.get(), c, mco[c]);
    s->t[c] = t[c];
    s->dt[c] = dt[c];
  }
  std::memcpy(s->U.data(), u, s->U.size() * sizeof(double));
  s->implicit.updates = counts[0];
  s->implicit.retries = counts[1];
  const int status = solver2d_run_implicit(s.get(), tend, static_cast<long>(maxsteps));
  std::memcpy(u, s->U.data(), s->U.size() * sizeof(double));
  for (int c = 0; c < ncase; ++c) {
    t[c] = s->t[c];
    dt[c] = s->dt[c];
  }
  *nsteps = s->nsteps;
  counts[0] = s->implicit.updates;
  counts[1] = s->implicit.retries;
  defects[0] = s->implicit.initial_defect;
  defects[1] = s->implicit.final_defect;
  defects[2] = s->implicit.tolerance;
  return status;
}

// cut2d_build_polygon (src/cutcell2d.cpp:247-503) on the grid of solver2d_init.  cell: 14 arrays of
// mx * my (vol, cx, cy, flag, wax, way, wcx, wcy, nbrcount, smallpos, then 4 unused); xface: axl,
// axy; yface: ayl, ayx; small: sidx, scnt, svhat, sxhat, syhat (capacity each) and smem (25 x
// capacity); counts: nsmall, nbad_nbhd, nsolid, ncut, walls.  walls: capacity_walls records.
int ref_grid_cut_build(int nx, int ny, const double* box, const ugrid_params* p, const int* stretch_type,
                       const double* stretch_param, const double* px, const double* py, int n, const int* patches,
                       double* const* cell, double* const* xface, double* const* yface, double* const* small,
                       double* smem, int capacity, int64_t* counts, amr_wall* walls, int capacity_walls) {
  auto s = ref_grid_solver(nx, ny, box, 1, p, stretch_type, stretch_param);
  Cut2D cut;
  const int status = cut2d_build_polygon(s.get(), px, py, n, &cut, patches);
  if (status) return status;
  const Array* cells[] = {&cut.vol, &cut.cx, &cut.cy, &cut.flag, &cut.wax, &cut.way, &cut.wcx, &cut.wcy,
                          &cut.nbrcount, &cut.smallpos};
  for (int k = 0; k < 10; ++k) std::memcpy(cell[k], cells[k]->data(), cells[k]->size() * sizeof(double));
  std::memcpy(xface[0], cut.axl.data(), cut.axl.size() * sizeof(double));
  std::memcpy(xface[1], cut.axy.data(), cut.axy.size() * sizeof(double));
  std::memcpy(yface[0], cut.ayl.data(), cut.ayl.size() * sizeof(double));
  std::memcpy(yface[1], cut.ayx.data(), cut.ayx.size() * sizeof(double));
  if (cut.nsmall > capacity || static_cast<int>(cut.wall_fragments.size()) > capacity_walls) return 9;
  const Array* smalls[] = {&cut.sidx, &cut.scnt, &cut.svhat, &cut.sxhat, &cut.syhat};
  for (int k = 0; k < 5; ++k) std::memcpy(small[k], smalls[k]->data(), smalls[k]->size() * sizeof(double));
  std::memcpy(smem, cut.smem.data(), cut.smem.size() * sizeof(double));
  static_assert(sizeof(amr_wall) == sizeof(WallFragment2D), "wall record layout");
  std::memcpy(walls, cut.wall_fragments.data(), cut.wall_fragments.size() * sizeof(WallFragment2D));
  counts[0] = cut.nsmall;
  counts[1] = cut.nbad_nbhd;
  counts[2] = cut.nsolid;
  counts[3] = cut.ncut;
  counts[4] = static_cast<int64_t>(cut.wall_fragments.size());
  return 0;
}

// The residual and solver2d_run (SSP-RK3 or backward Euler, by params->time_integrator) on the
// grid with the cut geometry of the polygon (cut2d_build_polygon).  counts: updates, retries
// (in/out); defects: initial, final, tolerance.
int ref_grid_cut_rhs(int nx, int ny, const double* box, int ncase, const ugrid_params* p, const int* stretch_type,
                     const double* stretch_param, const double* mco, const double* px, const double* py, int n,
                     const double* u, double* r) {
  auto s = ref_grid_solver(nx, ny, box, ncase, p, stretch_type, stretch_param);
  Cut2D cut;
  if (cut2d_build_polygon(s.get(), px, py, n, &cut, nullptr)) return 1;
  solver2d_set_cut(s.get(), &cut);
  if (mco)
    for (int c = 0; c < ncase; ++c) solver2d_set_mco(s.get(), c, mco[c]);
  std::memcpy(s->U.data(), u, s->U.size() * sizeof(double));
  solver2d_apply_bc(s.get(), s->U.data());
  solver2d_residual(s.get(), s->U.data(), s->R.data(), s->W.data(), s->SX.data(), s->SY.data(), s->FX.data(),
                    s->FY.data());
  std::memcpy(r, s->R.data(), s->R.size() * sizeof(double));
  return 0;
}

int ref_grid_cut_run(int nx, int ny, const double* box, int ncase, const ugrid_params* p, const int* stretch_type,
                     const double* stretch_param, const double* mco, const double* px, const double* py, int n,
                     double* u, double* t, double* dt, double tend, int64_t maxsteps, int64_t* nsteps,
                     int64_t* nreject, int64_t* counts, double* defects) {
  auto s = ref_grid_solver(nx, ny, box, ncase, p, stretch_type, stretch_param);
  Cut2D cut;
  if (cut2d_build_polygon(s.get(), px, py, n, &cut, nullptr)) return 9;
  solver2d_set_cut(s.get(), &cut);
  for (int c = 0; c < ncase; ++c) {
    if (mco) solver2d_set_mco(s.get(), c, mco[c]);
    s->t[c] = t[c];
    s->dt[c] = dt[c];
  }
  std::memcpy(s->U.data(), u, s->U.size() * sizeof(double));
  const int status = solver2d_run(s.get(), tend, static_cast<long>(maxsteps));
  std::memcpy(u, s->U.data(), s->U.size() * sizeof(double));
  for (int c = 0; c < ncase; ++c) {
    t[c] = s->t[c];
    dt[c] = s->dt[c];
  }
  *nsteps = s->nsteps;
  *nreject = s->nreject;
  counts[0] = s->implicit.updates;
  counts[1] = s->implicit.retries;
  defects[0] = s->implicit.initial_defect;
  defects[1] = s->implicit.final_defect;
  defects[2] = s->implicit.tolerance;
  return status;
}

// The same list as ugrid_solver2d_layout (include/ugrid_hip.h), measured on the reference's types.
int ref_solver2d_layout(int64_t* out, int32_t capacity) {
  const int64_t v[] = {
      sizeof(Solver2D), sizeof(Grid2D), sizeof(Array), sizeof(Params2D), sizeof(ImplicitStats2D),
      offsetof(Solver2D, grid), offsetof(Grid2D, nx), offsetof(Grid2D, ng), offsetof(Grid2D, mx), offsetof(Grid2D, my),
      offsetof(Grid2D, xmin), offsetof(Grid2D, xf), offsetof(Grid2D, yf), offsetof(Grid2D, xc), offsetof(Grid2D, dx),
      offsetof(Grid2D, yc), offsetof(Grid2D, dy), offsetof(Solver2D, par), offsetof(Solver2D, ncase),
      offsetof(Solver2D, U), offsetof(Solver2D, U1), offsetof(Solver2D, U2), offsetof(Solver2D, R),
      offsetof(Solver2D, W), offsetof(Solver2D, SX), offsetof(Solver2D, SY), offsetof(Solver2D, FX),
      offsetof(Solver2D, FY), offsetof(Solver2D, t), offsetof(Solver2D, dt), offsetof(Solver2D, mco),
      offsetof(Solver2D, nsteps), offsetof(Solver2D, nreject), offsetof(Solver2D, implicit), offsetof(Solver2D, ib),
      offsetof(Solver2D, cut), sizeof(Cut2D), offsetof(Cut2D, nsmall), offsetof(Cut2D, nbad_nbhd),
      offsetof(Cut2D, vol), offsetof(Cut2D, flag), offsetof(Cut2D, axl), offsetof(Cut2D, ayl), offsetof(Cut2D, wax),
      offsetof(Cut2D, way), offsetof(Cut2D, nbrcount), offsetof(Cut2D, smallpos), offsetof(Cut2D, scnt),
      offsetof(Cut2D, smem), offsetof(Cut2D, svhat)};
  const int32_t n = static_cast<int32_t>(sizeof(v) / sizeof(v[0]));
  for (int32_t k = 0; k < n && k < capacity; ++k) out[k] = v[k];
  return n;
}

// kind 0 HLLC, 1 AUSM+-up; states (rho, normal velocity, tangential velocity, p).
void ref_flux(int kind, const double* L, const double* R, double gamma, double mco, double* F) {
  if (kind == 1)
    flux_ausmpup_t(L[0], L[1], L[2], L[3], R[0], R[1], R[2], R[3], gamma, mco, F);
  else
    flux_hllc_t(L[0], L[1], L[2], L[3], R[0], R[1], R[2], R[3], gamma, F);
}

}  // extern "C"
