// viscous_prep.cpp -- fixed viscous tables of a built mesh (see mesher.h).
//
//   weighted least-squares velocity-gradient terms per cell ..... src/static_amr2d_viscous.cpp:44-55
//   wall donor: nearest visible cell within four face rings ...... src/static_amr2d_viscous.cpp:16-32,56-66
//   exterior face samples and fixed face geometry ................ src/static_amr2d_viscous.cpp:67-74
//   SA-neg geometric wall distance ............................... src/static_amr2d_sa.cpp:5-15
//   transport / boundary validity ................................ src/static_amr2d_viscous.h:148-156,
//                                                                  src/static_amr2d_boundary.h:32-47
// The tables are inputs of the time step and are compared bit for bit with the reference's.
#include <algorithm>
#include <cmath>
#include <limits>
#include <queue>

#include "mesher.h"

namespace cfd {
namespace hipamr {
namespace {

int failed(std::string* error, const std::string& message) {
  if (error) *error = message;
  return 1;
}
int side_kind(const StaticAmrMesh2D& m, const Params2D& p, int side) {
  return m.boundaries[side].kind < 0 ? p.bc[side] : m.boundaries[side].kind;
}
double cross(double x, double y, double a, double b) { return x * b - y * a; }

// Even-odd test against the stored body polygon (its own rounding, src/static_amr2d_viscous.cpp:16-22).
bool inside_body(const StaticAmrMesh2D& m, double x, double y) {
  bool inside = false;
  const size_t n = m.body_x.size();
  for (size_t i = 0, j = n ? n - 1 : 0; i < n; j = i++)
    if ((m.body_y[i] > y) != (m.body_y[j] > y)) {
      const double hit = (m.body_x[j] - m.body_x[i]) * (y - m.body_y[i]) / (m.body_y[j] - m.body_y[i]) + m.body_x[i];
      if (x < hit) inside = !inside;
    }
  return inside;
}
// Segment (x, y) -> cell centroid does not pass through the body.
bool sees(const StaticAmrMesh2D& m, double x, double y, const AmrCell2D& cell) {
  if (inside_body(m, cell.cx, cell.cy)) return false;
  const double dx = cell.cx - x, dy = cell.cy - y;
  const size_t nb = m.body_x.size();
  for (size_t i = 0; i < nb; ++i) {
    const size_t j = (i + 1) % nb;
    const double ex = m.body_x[j] - m.body_x[i], ey = m.body_y[j] - m.body_y[i];
    const double det = cross(dx, dy, ex, ey);
    if (std::fabs(det) <= 1e-14 * std::hypot(dx, dy) * std::hypot(ex, ey)) continue;
    const double ax = m.body_x[i] - x, ay = m.body_y[i] - y;
    const double t = cross(ax, ay, ex, ey) / det, u = cross(ax, ay, dx, dy) / det;
    if (t > 1e-10 && t <= 1 + 1e-12 && u >= -1e-12 && u <= 1 + 1e-12) return false;
  }
  return true;
}

int prepare_sa_distance(StaticAmrMesh2D& m, const Params2D& p, std::string* error) {
  m.viscous.wall_distance.clear();
  if (!m.transport.sa_model) return 0;
  if (!m.options.agglomerate_small || m.small_cells || m.transport.stabilization != AMR_VISCOUS_VOLUME)
    return failed(error, "SA-neg requires agglomerated true-volume mesh without SRD");
  const size_t nb = m.body_x.size();
  for (const AmrCell2D& c : m.cells) {
    double d = INFINITY;
    for (size_t j = 0; j < nb; ++j) {
      const size_t k = (j + 1) % nb;
      const double ax = m.body_x[j], ay = m.body_y[j], ex = m.body_x[k] - ax, ey = m.body_y[k] - ay;
      const double t = std::max(0., std::min(1., ((c.cx - ax) * ex + (c.cy - ay) * ey) / (ex * ex + ey * ey)));
      d = std::min(d, std::hypot(c.cx - ax - t * ex, c.cy - ay - t * ey));
    }
    for (int side = 0; side < 4; ++side)
      if (side_kind(m, p, side) == AMR_BC_NOSLIP) {
        const double v = side < 2 ? c.cx : c.cy;
        const double wall = side == 0 ? m.options.xmin : side == 1 ? m.options.xmax : side == 2 ? m.options.ymin : m.options.ymax;
        d = std::min(d, std::fabs(v - wall));
      }
    if (!(d > 0)) return failed(error, "SA geometric wall distance is nonpositive");
    m.viscous.wall_distance.push_back(d);
  }
  return 0;
}

}  // namespace

bool transport_valid(const AmrTransport2D& t) {
  if (t.model == AMR_MODEL_EULER) return t.sa_model == 0;
  return (t.sa_model == 0 || (t.sa_model == 1 && std::isfinite(t.sa_nu_inf) && t.sa_nu_inf >= 0 &&
                              std::isfinite(t.sa_cfl) && t.sa_cfl > 0 && t.sa_cfl <= .1)) &&
         t.model == AMR_MODEL_LAMINAR && (t.law == AMR_MU_CONSTANT || t.law == AMR_MU_SUTHERLAND) &&
         std::isfinite(t.mu_ref) && t.mu_ref > 0 && std::isfinite(t.gas_constant) && t.gas_constant > 0 &&
         std::isfinite(t.temperature_ref) && t.temperature_ref > 0 && std::isfinite(t.sutherland) &&
         t.sutherland >= 0 && std::isfinite(t.viscous_cfl) && t.viscous_cfl > 0 && t.viscous_cfl <= .2 &&
         (t.stabilization == AMR_VISCOUS_VOLUME || t.stabilization == AMR_VISCOUS_SRD) &&
         std::isfinite(t.wall_probe_min_h) && t.wall_probe_min_h >= 0 && t.wall_probe_min_h <= 2;
}

static bool primitive_ok(const double* w) {
  return std::isfinite(w[0]) && w[0] > 0 && std::isfinite(w[1]) && std::isfinite(w[2]) && std::isfinite(w[3]) && w[3] > 0;
}

bool boundary_valid(const AmrBoundary2D& b, int legacy, int side) {
  if (b.kind < AMR_BC_INHERIT) return false;
  const int kind = b.kind < 0 ? legacy : b.kind;
  if (kind == AMR_BC_TRANSMISSIVE || kind == AMR_BC_REFLECT) return true;
  if (kind == AMR_BC_NOSLIP)
    return std::isfinite(b.wall_velocity[0]) && std::isfinite(b.wall_velocity[1]) && b.wall_velocity[side / 2] == 0;
  if (kind < AMR_BC_FARFIELD || kind > AMR_BC_PRESSURE_OUTLET || !primitive_ok(b.reference)) return false;
  if (kind == AMR_BC_PRESSURE_OUTLET) return std::isfinite(b.outlet_pressure) && b.outlet_pressure > 0;
  if (kind == AMR_BC_TOTAL_INLET) {
    const double d2 = b.direction[0] * b.direction[0] + b.direction[1] * b.direction[1];
    const double dn = (side % 2 ? 1.0 : -1.0) * b.direction[side / 2];
    return std::isfinite(d2) && std::fabs(d2 - 1) < 1e-12 && dn < 0 && std::isfinite(b.total_pressure) &&
           b.total_pressure > 0 && std::isfinite(b.total_temperature) && b.total_temperature > 0 &&
           std::isfinite(b.gas_constant) && b.gas_constant > 0;
  }
  return true;
}

bool viscous_ready(const StaticAmrMesh2D& m) {
  if (m.transport.model == AMR_MODEL_EULER)
    for (const AmrBoundary2D& b : m.boundaries)
      if (b.kind == AMR_BC_NOSLIP) return false;
  const AmrViscousTable2D& v = m.viscous;
  return (!m.transport.sa_model || v.wall_distance.size() == m.cells.size()) && transport_valid(m.transport) &&
         (m.transport.model == AMR_MODEL_EULER ||
          (v.offsets.size() == m.cells.size() + 1 && v.wall_samples.size() == m.walls.size() &&
           v.face_samples.size() == m.faces.size() && v.face_geometry.size() == m.faces.size()));
}

int prepare_viscous(StaticAmrMesh2D* mesh, const Params2D& p, std::string* error) {
  if (!mesh) return failed(error, "null viscous mesh");
  StaticAmrMesh2D& m = *mesh;
  m.viscous = AmrViscousTable2D{};
  if (!transport_valid(m.transport)) return failed(error, "invalid laminar transport parameters");
  if (m.transport.model == AMR_MODEL_EULER)
    return viscous_ready(m) ? 0 : failed(error, "no-slip wall requires laminar model");
  const size_t n = m.cells.size();
  AmrViscousTable2D& table = m.viscous;
  table.offsets.push_back(0);
  // Least-squares weights: g = sum_s w_s (q_s - q_k), 1/d^2-weighted normal equations.
  struct Sample {
    int source, side;
    double dx, dy;
  };
  for (size_t k = 0; k < n; ++k) {
    const AmrCell2D& c = m.cells[k];
    std::vector<Sample> samples;
    for (size_t pos = m.face_offsets[k]; pos < m.face_offsets[k + 1]; ++pos) {
      const AmrFace2D& f = m.faces[m.cell_faces[pos]];
      const int other = f.left == static_cast<int>(k) ? f.right : f.left;
      if (other >= 0) {
        samples.push_back({other, -1, m.cells[other].cx - c.cx, m.cells[other].cy - c.cy});
      } else if (side_kind(m, p, f.boundary) == AMR_BC_NOSLIP) {
        const double x = f.axis == 0 ? f.coordinate : f.centroid, y = f.axis == 1 ? f.coordinate : f.centroid;
        samples.push_back({-2, f.boundary, x - c.cx, y - c.cy});
      }
    }
    for (size_t pos = c.wall_begin; pos < c.wall_begin + c.wall_count; ++pos) {
      const WallFragment2D& w = m.walls[pos];
      samples.push_back({-1, -1, .5 * (w.x0 + w.x1) - c.cx, .5 * (w.y0 + w.y1) - c.cy});
    }
    double xx = 0, xy = 0, yy = 0;
    for (const Sample& s : samples) {
      const double d2 = s.dx * s.dx + s.dy * s.dy;
      if (!(d2 > 0)) return failed(error, "zero gradient sample distance");
      xx += s.dx * s.dx / d2;
      xy += s.dx * s.dy / d2;
      yy += s.dy * s.dy / d2;
    }
    const double det = xx * yy - xy * xy;
    if (!(det > 1e-12 * (xx + yy) * (xx + yy)))
      return failed(error, "singular/ill-conditioned velocity gradient stencil at cell " + std::to_string(k));
    for (const Sample& s : samples) {
      const double d2 = s.dx * s.dx + s.dy * s.dy;
      table.terms.push_back({s.source, s.side, (yy * s.dx - xy * s.dy) / (det * d2), (xx * s.dy - xy * s.dx) / (det * d2)});
    }
    table.offsets.push_back(table.terms.size());
  }
  // Wall donor: breadth-first over at most four face rings, nearest visible fluid-side centroid.
  for (const WallFragment2D& w : m.walls) {
    const double x = .5 * (w.x0 + w.x1), y = .5 * (w.y0 + w.y1);
    const AmrCell2D& own = m.cells[w.owner_cell];
    const double h = std::min(own.x1 - own.x0, own.y1 - own.y0);
    std::queue<std::pair<int, int>> todo;
    std::vector<bool> visited(n, false);
    todo.push({static_cast<int>(w.owner_cell), 0});
    visited[w.owner_cell] = true;
    int donor = -1;
    double best = INFINITY, distance = 0;
    while (!todo.empty()) {
      const auto [k, depth] = todo.front();
      todo.pop();
      const AmrCell2D& c = m.cells[k];
      const double dx = x - c.cx, dy = y - c.cy, dn = dx * w.nx + dy * w.ny, d2 = dx * dx + dy * dy;
      if (dn > 1e-12 * h && dn >= m.transport.wall_probe_min_h * h && d2 < best && sees(m, x, y, c)) {
        donor = k;
        distance = dn;
        best = d2;
      }
      if (depth == 4) continue;
      for (size_t pos = m.face_offsets[k]; pos < m.face_offsets[k + 1]; ++pos) {
        const AmrFace2D& f = m.faces[m.cell_faces[pos]];
        const int other = f.left == k ? f.right : f.left;
        if (other >= 0 && !visited[other]) {
          visited[other] = true;
          todo.push({other, depth + 1});
        }
      }
    }
    if (donor < 0) return failed(error, "no visible fluid-side wall donor within four face rings");
    table.wall_samples.push_back({donor, distance});
  }
  // Exterior faces sample their owner; interior faces keep the centroid-to-centroid geometry.
  for (const AmrFace2D& f : m.faces) {
    AmrWallSample2D sample{-1, 0};
    if (f.boundary >= 0) {
      const int owner = f.left >= 0 ? f.left : f.right;
      const AmrCell2D& c = m.cells[owner];
      const double sign = f.boundary % 2 ? 1 : -1;
      sample = {owner, sign * (f.coordinate - (f.axis == 0 ? c.cx : c.cy))};
      if (!(sample.distance > 0)) return failed(error, "invalid exterior face distance");
    }
    table.face_samples.push_back(sample);
    AmrViscousFaceGeometry2D geo{};
    const double x = f.axis == 0 ? f.coordinate : f.centroid, y = f.axis == 1 ? f.coordinate : f.centroid;
    if (f.boundary < 0) {
      const AmrCell2D& l = m.cells[f.left];
      const AmrCell2D& r = m.cells[f.right];
      geo = {r.cx - l.cx, r.cy - l.cy, x - .5 * (l.cx + r.cx), y - .5 * (l.cy + r.cy), std::hypot(r.cx - l.cx, r.cy - l.cy)};
    } else {
      const AmrCell2D& c = m.cells[sample.donor];
      geo = {0, 0, x - c.cx, y - c.cy, sample.distance};
    }
    if (!(geo.distance > 0) || !std::isfinite(geo.distance)) return failed(error, "invalid fixed viscous face geometry");
    table.face_geometry.push_back(geo);
  }
  return prepare_sa_distance(m, p, error);
}

}  // namespace hipamr
}  // namespace cfd
