// cutcell.cpp -- rectangle/polygon cut geometry (see cutcell.h).
//
// The arithmetic follows the reference primitives so the mesh is bit-identical:
//   crossing of an edge with a grid line ............ src/cutcell2d.cpp:27-37
//   wall lying on a grid line ........................ src/cutcell2d.cpp:39-45
//   aperture length/centroid of a grid segment ....... src/cutcell2d.cpp:47-76, 117-122
//   Liang-Barsky clipping of an edge ................. src/cutcell2d.cpp:78-102
//   fluid-side edge normal ........................... src/cutcell2d.cpp:104-113
//   rectangle-to-polygon distance .................... src/cutcell2d.cpp:124-140
//   one-region metrics by boundary integrals ......... src/cutcell2d.cpp:142-194
//   open intervals / multi-region loop tracing ....... src/cutcell2d.cpp:196-245
//   point-in-polygon and segment distance ............ src/sdf.h:15-42
#include "cutcell.h"

#include <algorithm>
#include <cmath>
#include <limits>

namespace cfd {
namespace hipamr {

double segment_distance2(double x, double y, double ax, double ay, double bx, double by) {
  const double ex = bx - ax, ey = by - ay;
  const double e2 = ex * ex + ey * ey;
  double t = 0.0;
  if (e2 > 0.0) {
    t = ((x - ax) * ex + (y - ay) * ey) / e2;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  }
  const double qx = ax + t * ex - x, qy = ay + t * ey - y;
  return qx * qx + qy * qy;
}

// Even-odd ray cast towards +x; edges visited as (n-1 -> 0), (0 -> 1), ...
bool inside_polygon(const Polygon& b, double x, double y) {
  bool inside = false;
  for (int i = 0, j = b.n - 1; i < b.n; j = i++) {
    if ((b.y[i] > y) != (b.y[j] > y)) {
      const double xc = b.x[j] + (y - b.y[j]) * (b.x[i] - b.x[j]) / (b.y[i] - b.y[j]);
      if (x < xc) inside = !inside;
    }
  }
  return inside;
}

namespace {

// Coordinates of vertex k across (c) and along (t) a grid line of the given axis.
inline double across(const Polygon& b, int axis, int k) { return axis == 0 ? b.x[k] : b.y[k]; }
inline double along(const Polygon& b, int axis, int k) { return axis == 0 ? b.y[k] : b.x[k]; }

// Crossing of edge a->b with the line {across = c}, half-open convention (a>=c) != (b>=c).
inline bool edge_crosses(const Polygon& b, int axis, int a, int e, double c, double* t) {
  const double ca = across(b, axis, a), ce = across(b, axis, e);
  if ((ca >= c) == (ce >= c)) return false;
  const double ta = along(b, axis, a), te = along(b, axis, e);
  *t = ta + (c - ca) * (te - ta) / (ce - ca);
  return true;
}

// True when a polygon edge lies on the line {across = c} and covers the point t.
bool on_line_wall(const Polygon& b, int axis, double c, double t) {
  for (int a = 0; a < b.n; ++a) {
    const int e = (a + 1) % b.n;
    if (across(b, axis, a) == c && across(b, axis, e) == c) {
      const double ta = along(b, axis, a), te = along(b, axis, e);
      if (t >= std::min(ta, te) && t <= std::max(ta, te)) return true;
    }
  }
  return false;
}

inline bool solid_on_line(const Polygon& b, int axis, double c, double t) {
  return axis == 0 ? inside_polygon(b, c, t) : inside_polygon(b, t, c);
}

}  // namespace

bool clip_to_rect(double* ax, double* ay, double* bx, double* by, double x0, double x1, double y0,
                  double y1) {
  double t0 = 0.0, t1 = 1.0;
  const double dx = *bx - *ax, dy = *by - *ay;
  const double p[4] = {-dx, dx, -dy, dy};
  const double q[4] = {*ax - x0, x1 - *ax, *ay - y0, y1 - *ay};
  for (int k = 0; k < 4; ++k) {
    if (p[k] == 0.0) {
      if (q[k] < 0.0) return false;
      continue;
    }
    const double r = q[k] / p[k];
    if (p[k] < 0.0) {
      if (r > t1) return false;
      if (r > t0) t0 = r;
    } else {
      if (r < t0) return false;
      if (r < t1) t1 = r;
    }
  }
  const double sx = *ax + t0 * dx, sy = *ay + t0 * dy, ex = *ax + t1 * dx, ey = *ay + t1 * dy;
  *ax = sx;
  *ay = sy;
  *bx = ex;
  *by = ey;
  return (t1 - t0) > 1e-14;
}

void fluid_normal(const Polygon& b, int a, double mx, double my, double eps, double* nx,
                  double* ny) {
  const int e = (a + 1) % b.n;
  const double tx = b.x[e] - b.x[a], ty = b.y[e] - b.y[a];
  const double len = std::sqrt(tx * tx + ty * ty);
  double ux = ty / len, uy = -tx / len;
  if (inside_polygon(b, mx + eps * ux, my + eps * uy)) {
    ux = -ux;
    uy = -uy;
  }
  *nx = ux;
  *ny = uy;
}

void open_aperture(const Polygon& b, int axis, double c, double lo, double hi, double* length,
                   double* centroid) {
  open_aperture_raw(b, axis, c, lo, hi, length, centroid);
  // Snap near-full and near-empty apertures (src/cutcell2d.cpp:120-121).
  const double len = *length;
  if (std::fabs(len - (hi - lo)) < 1e-13 * (hi - lo)) *length = hi - lo;
  if (len < 1e-13 * (hi - lo)) *length = 0.0;
}

void open_aperture_raw(const Polygon& b, int axis, double c, double lo, double hi, double* length,
                       double* centroid) {
  std::vector<double> cuts;
  cuts.push_back(lo);
  for (int i = 0, j = b.n - 1; i < b.n; j = i++) {
    if (across(b, axis, j) == c && across(b, axis, i) == c) {
      for (double t : {along(b, axis, j), along(b, axis, i)})
        if (t > lo && t < hi) cuts.push_back(t);
    }
    double t;
    if (edge_crosses(b, axis, j, i, c, &t) && t > lo && t < hi) cuts.push_back(t);
  }
  cuts.push_back(hi);
  std::sort(cuts.begin(), cuts.end());
  double len = 0.0, moment = 0.0;
  for (size_t k = 0; k + 1 < cuts.size(); ++k) {
    const double a = cuts[k], e = cuts[k + 1];
    if (e - a <= 0.0) continue;
    const double mid = 0.5 * (a + e);
    if (!solid_on_line(b, axis, c, mid) && !on_line_wall(b, axis, c, mid)) {
      len += e - a;
      moment += 0.5 * (e * e - a * a);
    }
  }
  *centroid = (len > 0.0) ? moment / len : 0.5 * (lo + hi);
  *length = len;
}

std::vector<Point> open_intervals(const Polygon& b, int axis, double c, double lo, double hi) {
  std::vector<double> cuts = {lo, hi};
  for (int a = 0; a < b.n; ++a) {
    const int e = (a + 1) % b.n;
    if (across(b, axis, a) == c && across(b, axis, e) == c) {
      for (double t : {along(b, axis, a), along(b, axis, e)})
        if (t > lo && t < hi) cuts.push_back(t);
    }
    double t;
    if (edge_crosses(b, axis, a, e, c, &t) && t > lo && t < hi) cuts.push_back(t);
  }
  std::sort(cuts.begin(), cuts.end());
  std::vector<Point> result;
  for (size_t i = 0; i + 1 < cuts.size(); ++i) {
    if (!(cuts[i + 1] > cuts[i])) continue;
    const double mid = .5 * (cuts[i] + cuts[i + 1]);
    if (!on_line_wall(b, axis, c, mid) && !solid_on_line(b, axis, c, mid))
      result.push_back({cuts[i], cuts[i + 1]});
  }
  return result;
}

double rect_wall_distance(const Polygon& b, double x0, double x1, double y0, double y1) {
  double d2 = std::numeric_limits<double>::infinity();
  for (int a = 0; a < b.n; ++a) {
    const int e = (a + 1) % b.n;
    double ax = b.x[a], ay = b.y[a], ex = b.x[e], ey = b.y[e];
    if (clip_to_rect(&ax, &ay, &ex, &ey, x0, x1, y0, y1)) return 0.0;
    for (double x : {x0, x1})
      for (double y : {y0, y1}) d2 = std::min(d2, segment_distance2(x, y, b.x[a], b.y[a], b.x[e], b.y[e]));
    for (int k : {a, e}) {
      const double gx = std::max({x0 - b.x[k], 0.0, b.x[k] - x1});
      const double gy = std::max({y0 - b.y[k], 0.0, b.y[k] - y1});
      d2 = std::min(d2, gx * gx + gy * gy);
    }
  }
  return std::sqrt(d2);
}

int rect_cut(const Polygon& b, double x0, double x1, double y0, double y1, CutRegion* r) {
  r->walls.clear();
  r->boundary.clear();
  r->volume = (x1 - x0) * (y1 - y0);
  r->cx = 0.5 * (x0 + x1);
  r->cy = 0.5 * (y0 + y1);
  if (b.n == 0) return 0;
  const double h = std::min(x1 - x0, y1 - y0);
  double area = 0, mx = 0, my = 0;
  std::vector<Point> ends;  // distinct points where the wall meets the rectangle outline
  for (int a = 0; a < b.n; ++a) {
    const int e = (a + 1) % b.n;
    double ax = b.x[a], ay = b.y[a], ex = b.x[e], ey = b.y[e];
    if (!clip_to_rect(&ax, &ay, &ex, &ey, x0, x1, y0, y1)) continue;
    const double length = std::hypot(ex - ax, ey - ay);
    double nfx, nfy;
    fluid_normal(b, a, 0.5 * (b.x[a] + b.x[e]), 0.5 * (b.y[a] + b.y[e]), 1e-7 * h, &nfx, &nfy);
    // Orient the clipped piece so the fluid lies on its left.
    if ((ex - ax) * nfy - (ey - ay) * nfx < 0) {
      std::swap(ax, ex);
      std::swap(ay, ey);
    }
    // A piece lying on the rectangle outline belongs only to the leaf on its fluid side.
    if ((ax == ex && ((ax == x0 && nfx < 0) || (ax == x1 && nfx > 0))) ||
        (ay == ey && ((ay == y0 && nfy < 0) || (ay == y1 && nfy > 0))))
      continue;
    // Moments in leaf-local coordinates (no cancellation of large global values).
    const double lax = ax - x0, lex = ex - x0, lay = ay - y0, ley = ey - y0;
    area += 0.5 * (lax + lex) * (ley - lay);
    mx += (lax * lax + lax * lex + lex * lex) * (ley - lay) / 6;
    my -= (lay * lay + lay * ley + ley * ley) * (lex - lax) / 6;
    WallFragment2D w{};
    w.owner_cell = 0;
    w.source_edge_id = a;
    w.body_id = 0;
    w.patch_id = b.patch ? b.patch[a] : 0;
    w.x0 = ax;
    w.y0 = ay;
    w.x1 = ex;
    w.y1 = ey;
    w.nx = -nfx;
    w.ny = -nfy;
    w.length = length;
    r->walls.push_back(w);
    for (const Point& p : {Point{ax, ay}, Point{ex, ey}}) {
      const double tol = 1e-11 * h;
      if (std::min({std::fabs(p[0] - x0), std::fabs(p[0] - x1), std::fabs(p[1] - y0),
                    std::fabs(p[1] - y1)}) > tol)
        continue;
      bool seen = false;
      for (const Point& q : ends)
        if (std::hypot(p[0] - q[0], p[1] - q[1]) < tol) seen = true;
      if (!seen) ends.push_back(p);
    }
  }
  if (r->walls.empty()) {
    if (inside_polygon(b, r->cx, r->cy)) r->volume = 0;
    return 0;
  }
  if (ends.size() != 2) return 2;  // more than one wall arc: traced by rect_regions
  double west, east, south, north, c;
  open_aperture(b, 0, x0, y0, y1, &west, &c);
  open_aperture(b, 0, x1, y0, y1, &east, &c);
  open_aperture(b, 1, y0, x0, x1, &south, &c);
  open_aperture(b, 1, y1, x0, x1, &north, &c);
  (void)west;
  (void)south;
  area += (x1 - x0) * east;
  mx += 0.5 * (x1 - x0) * (x1 - x0) * east;
  my += 0.5 * (y1 - y0) * (y1 - y0) * north;
  if (!std::isfinite(area) || area > r->volume * (1 + 1e-10) || area < -r->volume * 1e-10) return 3;
  if (area < 1e-13 * r->volume) {
    r->volume = 0;
    r->walls.clear();
    return 0;
  }
  r->volume = area;
  r->cx = x0 + mx / area;
  r->cy = y0 + my / area;
  return 0;
}

int rect_regions(const Polygon& b, double x0, double x1, double y0, double y1,
                 std::vector<CutRegion>* out) {
  out->clear();
  CutRegion whole;
  const int status = rect_cut(b, x0, x1, y0, y1, &whole);
  if (status == 0) {
    if (whole.volume > 0) out->push_back(std::move(whole));
    return 0;
  }
  if (status != 2) return status;
  // Several wall arcs: trace each oriented fluid loop as its own control volume.
  struct Edge {
    int from, to, wall;
    bool used;
  };
  std::vector<Edge> edges;
  std::vector<Point> nodes;
  const double h = std::min(x1 - x0, y1 - y0), tol = 1e-11 * h;
  const auto node = [&](double x, double y) {
    for (size_t k = 0; k < nodes.size(); ++k)
      if (std::hypot(nodes[k][0] - x, nodes[k][1] - y) <= tol) return static_cast<int>(k);
    nodes.push_back({x, y});
    return static_cast<int>(nodes.size() - 1);
  };
  const auto add_edge = [&](double ax, double ay, double ex, double ey, int wall) {
    const int a = node(ax, ay), e = node(ex, ey);
    if (a != e) edges.push_back({a, e, wall, false});
  };
  for (size_t q = 0; q < whole.walls.size(); ++q) {
    const auto& w = whole.walls[q];
    add_edge(w.x0, w.y0, w.x1, w.y1, static_cast<int>(q));
  }
  // Open pieces of the outline, oriented counter-clockwise around the leaf.
  for (int side = 0; side < 4; ++side) {
    const int axis = side < 2 ? 0 : 1;
    const double c = side == 0 ? x0 : side == 1 ? x1 : side == 2 ? y0 : y1;
    for (const Point& iv : open_intervals(b, axis, c, axis == 0 ? y0 : x0, axis == 0 ? y1 : x1)) {
      const double lo = iv[0], hi = iv[1];
      if (side == 0) add_edge(c, hi, c, lo, -1);
      if (side == 1) add_edge(c, lo, c, hi, -1);
      if (side == 2) add_edge(lo, c, hi, c, -1);
      if (side == 3) add_edge(hi, c, lo, c, -1);
    }
  }
  const double pi = std::acos(-1.);
  for (size_t start = 0; start < edges.size(); ++start) {
    if (edges[start].used) continue;
    CutRegion r{};
    size_t current = start;
    const int first = edges[start].from;
    long double twice_area = 0, mx = 0, my = 0;
    for (size_t count = 0;; ++count) {
      if (count > edges.size() || edges[current].used) return 4;
      Edge& e = edges[current];
      e.used = true;
      const Point a = nodes[e.from], z = nodes[e.to];
      r.boundary.push_back(a);
      const long double ax = a[0] - x0, ay = a[1] - y0, zx = z[0] - x0, zy = z[1] - y0;
      const long double cross = ax * zy - zx * ay;
      twice_area += cross;
      mx += (ax + zx) * cross;
      my += (ay + zy) * cross;
      if (e.wall >= 0) r.walls.push_back(whole.walls[e.wall]);
      if (e.to == first) break;
      // Next edge: the sharpest clockwise turn from the reversed incoming direction.
      const double reverse = std::atan2(a[1] - z[1], a[0] - z[0]);
      double best = INFINITY;
      size_t next = edges.size();
      for (size_t q = 0; q < edges.size(); ++q) {
        if (edges[q].used || edges[q].from != e.to) continue;
        const Point end = nodes[edges[q].to];
        double turn = reverse - std::atan2(end[1] - z[1], end[0] - z[0]);
        while (turn < 0) turn += 2 * pi;
        while (turn >= 2 * pi) turn -= 2 * pi;
        if (turn < best) {
          best = turn;
          next = q;
        }
      }
      if (next == edges.size()) return 4;
      current = next;
    }
    if (twice_area <= 0) return 5;  // a hole or reversed loop is never filled silently
    r.volume = static_cast<double>(.5L * twice_area);
    r.cx = x0 + static_cast<double>(mx / (3 * twice_area));
    r.cy = y0 + static_cast<double>(my / (3 * twice_area));
    if (!std::isfinite(r.volume) || r.volume > (x1 - x0) * (y1 - y0) * (1 + 1e-10)) return 5;
    out->push_back(std::move(r));
  }
  return out->empty() ? 5 : 0;
}

}  // namespace hipamr
}  // namespace cfd
