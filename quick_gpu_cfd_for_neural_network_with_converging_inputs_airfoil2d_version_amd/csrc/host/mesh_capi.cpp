// mesh_capi.cpp -- extern "C" entry points of the host mesher (amr_mesh_* in amr_hip.h).
#include <cmath>
#include <cstring>
#include <new>

#include "amr_hip.h"
#include "mesher.h"
#include "views.h"

using namespace cfd;
using namespace cfd::hipamr;

struct amr_mesh {
  StaticAmrMesh2D mesh;
};

extern "C" {

int amr_mesh_build(const amr_build_options* options, const double* px, const double* py, const int32_t* patches,
                   int32_t n_points, amr_mesh** out, char* err, size_t err_len) {
  if (!out || !options || n_points < 0 || (n_points > 0 && (!px || !py)) || options->n_boxes < 0 ||
      (options->n_boxes > 0 && !options->boxes))
    return report(err, err_len, "invalid mesh build arguments");
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
  for (int32_t b = 0; b < options->n_boxes; ++b) {
    const amr_box& box = options->boxes[b];
    o.boxes.push_back({box.x0, box.x1, box.y0, box.y1, box.level});
  }
  std::vector<double> x(px, px + n_points), y(py, py + n_points);
  std::vector<int> pt;
  if (patches) pt.assign(patches, patches + n_points);
  amr_mesh* m = new (std::nothrow) amr_mesh;
  if (!m) return report(err, err_len, "out of host memory");
  std::string error;
  if (build_mesh(o, x, y, pt, &m->mesh, &error)) {
    delete m;
    return report(err, err_len, error);
  }
  *out = m;
  return 0;
}

int amr_mesh_set_physics(amr_mesh* m, const amr_boundary* boundaries4, const amr_transport* transport) {
  if (!m) return 1;
  if (boundaries4) std::memcpy(static_cast<void*>(m->mesh.boundaries), boundaries4, 4 * sizeof(amr_boundary));
  if (transport) std::memcpy(static_cast<void*>(&m->mesh.transport), transport, sizeof(amr_transport));
  return 0;
}

int amr_mesh_prepare_viscous(amr_mesh* m, const amr_params* params, char* err, size_t err_len) {
  if (!m || !params) return report(err, err_len, "null viscous mesh");
  std::string error;
  if (prepare_viscous(&m->mesh, params_from_c(*params), &error)) return report(err, err_len, error);
  return 0;
}

int amr_mesh_get_view(const amr_mesh* m, amr_mesh_view* out) {
  if (!m || !out) return 1;
  fill_view(m->mesh, out);
  return 0;
}

int amr_mesh_get_info(const amr_mesh* m, amr_mesh_info* out) {
  if (!m || !out) return 1;
  const StaticAmrMesh2D& s = m->mesh;
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

int amr_mesh_initialize(const amr_mesh* m, const amr_params* params, const double* primitive4, double* u,
                        uint64_t u_len) {
  if (!m || !params || !primitive4 || !u) return 1;
  StaticAmrState2D state;
  const std::array<double, 4> w = {primitive4[0], primitive4[1], primitive4[2], primitive4[3]};
  if (initialize_state(m->mesh, params_from_c(*params), w, &state)) return 1;
  if (state.u.size() != u_len) return 1;
  std::memcpy(u, state.u.data(), u_len * sizeof(double));
  return 0;
}

int amr_mesh_copy_partition(const amr_mesh* m, int32_t* out, uint64_t rows) {
  if (!m || !out || rows != m->mesh.partition.size()) return 1;
  for (size_t k = 0; k < rows; ++k)
    for (int j = 0; j < 5; ++j) out[5 * k + j] = m->mesh.partition[k][j];
  return 0;
}

int amr_mesh_copy_leaf_components(const amr_mesh* m, uint64_t* out, uint64_t rows) {
  if (!m || !out || rows != m->mesh.leaf_components.size()) return 1;
  for (size_t k = 0; k < rows; ++k) {
    out[2 * k] = m->mesh.leaf_components[k][0];
    out[2 * k + 1] = m->mesh.leaf_components[k][1];
  }
  return 0;
}

int amr_mesh_copy_agglomeration(const amr_mesh* m, double* out, uint64_t rows) {
  if (!m || !out || rows != m->mesh.agglomeration.size()) return 1;
  for (size_t k = 0; k < rows; ++k)
    for (int j = 0; j < 9; ++j) out[9 * k + j] = m->mesh.agglomeration[k][j];
  return 0;
}

void amr_mesh_destroy(amr_mesh* m) { delete m; }

}  // extern "C"
