// views.h -- conversions between the C++ mesh/parameter types and the C ABI views.
#ifndef HIPAMR_VIEWS_H
#define HIPAMR_VIEWS_H

#include <cstring>
#include <string>

#include "amr2d_types.h"
#include "amr_hip.h"

namespace cfd {
namespace hipamr {

static_assert(sizeof(amr_cell) == sizeof(AmrCell2D), "amr_cell layout");
static_assert(sizeof(amr_face) == sizeof(AmrFace2D), "amr_face layout");
static_assert(sizeof(amr_wall) == sizeof(WallFragment2D), "amr_wall layout");
static_assert(sizeof(amr_grad_term) == sizeof(AmrGradientTerm2D), "amr_grad_term layout");
static_assert(sizeof(amr_wall_sample) == sizeof(AmrWallSample2D), "amr_wall_sample layout");
static_assert(sizeof(amr_face_geometry) == sizeof(AmrViscousFaceGeometry2D), "amr_face_geometry layout");
static_assert(sizeof(amr_boundary) == sizeof(AmrBoundary2D), "amr_boundary layout");
static_assert(sizeof(amr_transport) == sizeof(AmrTransport2D), "amr_transport layout");
static_assert(sizeof(uint64_t) == sizeof(size_t), "size_t width");

inline amr_params params_to_c(const Params2D& p) {
  amr_params c{};
  c.gamma = p.gamma;
  c.cfl = p.cfl;
  c.flux = p.flux;
  c.limiter = p.limiter;
  c.ausm_mco = p.ausm_mco;
  for (int k = 0; k < 4; ++k) c.bc[k] = p.bc[k];
  c.dt_fixed = p.dt_fixed;
  c.time_integrator = p.time_integrator;
  return c;
}

// Defaults of the reference's params2d_default (src/euler2d.cpp:13-37) for unused fields.
inline Params2D params_from_c(const amr_params& c) {
  Params2D p{};
  p.gamma = c.gamma;
  p.cfl = c.cfl;
  p.flux = c.flux;
  p.limiter = c.limiter;
  p.venkat_k = 5.0;
  p.ausm_mco = c.ausm_mco;
  for (int k = 0; k < 4; ++k) p.bc[k] = c.bc[k];
  p.dt_fixed = c.dt_fixed;
  p.time_integrator = c.time_integrator;
  p.pseudo_cfl = 0.4;
  p.implicit_maxiter = 2000;
  p.implicit_rtol = 1e-8;
  p.implicit_atol = 1e-11;
  return p;
}

inline void fill_view(const StaticAmrMesh2D& m, amr_mesh_view* v) {
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
}

inline int report(char* err, size_t len, const std::string& message) {
  if (err && len) {
    const size_t k = message.size() < len - 1 ? message.size() : len - 1;
    std::memcpy(err, message.data(), k);
    err[k] = 0;
  }
  return 1;
}

}  // namespace hipamr
}  // namespace cfd

#endif
