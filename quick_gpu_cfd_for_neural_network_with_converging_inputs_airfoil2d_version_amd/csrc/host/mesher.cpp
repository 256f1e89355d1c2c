// mesher.cpp -- static 2:1 quadtree cut-cell mesh (see mesher.h).
//
// The output arrays (cells, faces, walls, CSR, agglomeration, SRD groups) are a contract:
// they must equal the reference mesher's bit for bit, so the refinement order, the line
// grouping of std::map, std::sort on the same keys and the rounding of every geometric
// quantity follow src/static_amr2d.cpp:16-141,229-358 step by step:
//   leaf sides grouped per grid line and paired into subfaces ..... :20-52
//   leaf bounds / children ......................................... :54-66
//   SRD neighbourhoods (groups) .................................... :68-104
//   small-cell agglomeration ....................................... :106-141
//   refinement, 2:1 balance, cut regions, faces, CSR, closure ...... :229-358
#include "mesher.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <numeric>
#include <sstream>

#include "cutcell.h"

namespace cfd {
namespace hipamr {
namespace {

bool fail(std::string* error, const std::string& message) {
  if (error) *error = message;
  return false;
}

// A leaf side lying on one grid line: integer extent [lo, hi) along the line.
struct LineSide {
  int lo, hi, cell;
  bool low_side;  // the leaf lies below/left of the line (its +axis side is on the line)
};
// A subface between two leaves (-1 = domain outside), in integer finest-grid units.
struct Link {
  int left, right, axis, coordinate, lo, hi;
};

std::vector<Link> pair_sides(const std::vector<AmrCell2D>& leaves) {
  std::map<std::pair<int, int>, std::vector<LineSide>> lines;  // key (axis, coordinate)
  for (size_t k = 0; k < leaves.size(); ++k) {
    const AmrCell2D& c = leaves[k];
    const int id = static_cast<int>(k);
    lines[{0, c.ix}].push_back({c.iy, c.iy + c.span, id, false});
    lines[{0, c.ix + c.span}].push_back({c.iy, c.iy + c.span, id, true});
    lines[{1, c.iy}].push_back({c.ix, c.ix + c.span, id, false});
    lines[{1, c.iy + c.span}].push_back({c.ix, c.ix + c.span, id, true});
  }
  std::vector<Link> links;
  for (auto& line : lines) {
    const int axis = line.first.first, coordinate = line.first.second;
    std::vector<LineSide> below, above;
    for (const LineSide& s : line.second) (s.low_side ? below : above).push_back(s);
    const auto by_lo = [](const LineSide& a, const LineSide& b) { return a.lo < b.lo; };
    std::sort(below.begin(), below.end(), by_lo);
    std::sort(above.begin(), above.end(), by_lo);
    if (below.empty() || above.empty()) {  // domain boundary line
      for (const LineSide& s : line.second)
        links.push_back({s.low_side ? s.cell : -1, s.low_side ? -1 : s.cell, axis, coordinate, s.lo, s.hi});
      continue;
    }
    size_t i = 0, j = 0;  // merge the two tilings of the line into common subintervals
    while (i < below.size() && j < above.size()) {
      const LineSide& a = below[i];
      const LineSide& b = above[j];
      const int lo = std::max(a.lo, b.lo), hi = std::min(a.hi, b.hi);
      if (hi > lo) links.push_back({a.cell, b.cell, axis, coordinate, lo, hi});
      const int end = std::min(a.hi, b.hi);
      if (a.hi == end) ++i;
      if (b.hi == end) ++j;
    }
  }
  return links;
}

void set_bounds(AmrCell2D* c, const StaticAmrOptions2D& o) {
  const double dx = (o.xmax - o.xmin) / (o.nx * (1 << o.levels));
  const double dy = (o.ymax - o.ymin) / (o.ny * (1 << o.levels));
  c->x0 = o.xmin + c->ix * dx;
  c->x1 = o.xmin + (c->ix + c->span) * dx;
  c->y0 = o.ymin + c->iy * dy;
  c->y1 = o.ymin + (c->iy + c->span) * dy;
}

void split_leaf(const AmrCell2D& c, const StaticAmrOptions2D& o, std::vector<AmrCell2D>* out) {
  const int half = c.span / 2;
  for (int y = 0; y < 2; ++y)
    for (int x = 0; x < 2; ++x) {
      AmrCell2D child = {};
      child.ix = c.ix + x * half;
      child.iy = c.iy + y * half;
      child.span = half;
      child.level = c.level + 1;
      set_bounds(&child, o);
      out->push_back(child);
    }
}

// Depth-first refinement of one background cell (children in y-major order).
struct Refiner {
  const StaticAmrOptions2D& o;
  const Polygon& body;
  std::vector<AmrCell2D>& leaves;
  bool too_many = false;
  void visit(AmrCell2D c) {
    if (too_many) return;
    set_bounds(&c, o);
    bool refine = false;
    if (c.level < o.levels) {
      if (body.n > 0)
        refine = rect_wall_distance(body, c.x0, c.x1, c.y0, c.y1) <=
                 o.buffer * 0.5 * std::min(c.x1 - c.x0, c.y1 - c.y0);
      for (const AmrBox2D& box : o.boxes)
        if (c.level < box.level && c.x0 < box.x1 && c.x1 > box.x0 && c.y0 < box.y1 && c.y1 > box.y0)
          refine = true;
    }
    if (refine) {
      std::vector<AmrCell2D> sub;
      split_leaf(c, o, &sub);
      for (const AmrCell2D& child : sub) visit(child);
    } else {
      leaves.push_back(c);
      if (leaves.size() > o.max_cells) too_many = true;
    }
  }
};

// Cell -> face CSR in face order (src/static_amr2d.cpp:346-350).
void build_cell_faces(StaticAmrMesh2D* m) {
  const size_t n = m->cells.size();
  m->face_offsets.assign(n + 1, 0);
  for (const AmrFace2D& f : m->faces) {
    if (f.left >= 0) ++m->face_offsets[f.left + 1];
    if (f.right >= 0) ++m->face_offsets[f.right + 1];
  }
  for (size_t k = 0; k < n; ++k) m->face_offsets[k + 1] += m->face_offsets[k];
  m->cell_faces.resize(m->face_offsets.back());
  std::vector<size_t> cursor = m->face_offsets;
  for (size_t q = 0; q < m->faces.size(); ++q) {
    const AmrFace2D& f = m->faces[q];
    if (f.left >= 0) m->cell_faces[cursor[f.left]++] = q;
    if (f.right >= 0) m->cell_faces[cursor[f.right]++] = q;
  }
}

// Zero-degree SRD neighbourhoods: group g is cell g, grown breadth-first through faces until
// it holds half the cell's background rectangle, unless small cells are agglomerated.
bool build_groups(StaticAmrMesh2D* m, std::string* error) {
  const size_t n = m->cells.size();
  m->group_offsets = {0};
  m->group_members.clear();
  m->small_cells = 0;
  std::vector<int> mark(n, -1);
  for (size_t k = 0; k < n; ++k) {
    const AmrCell2D& cell = m->cells[k];
    const double target = 0.5 * (cell.x1 - cell.x0) * (cell.y1 - cell.y0);
    std::vector<size_t> group = {k};
    double volume = cell.volume;
    mark[k] = static_cast<int>(k);
    if (volume < target && !m->options.agglomerate_small) {
      ++m->small_cells;
      for (size_t pos = 0; pos < group.size() && volume < target; ++pos) {
        const size_t a = group[pos];
        for (size_t q = m->face_offsets[a]; q < m->face_offsets[a + 1] && volume < target; ++q) {
          const AmrFace2D& f = m->faces[m->cell_faces[q]];
          const int other = f.left == static_cast<int>(a) ? f.right : f.left;
          if (other < 0 || mark[other] == static_cast<int>(k)) continue;
          mark[other] = static_cast<int>(k);
          group.push_back(static_cast<size_t>(other));
          volume += m->cells[other].volume;
        }
      }
      if (volume < target) return fail(error, "small-cell neighborhood has insufficient connected fluid volume");
    }
    m->group_members.insert(m->group_members.end(), group.begin(), group.end());
    m->group_offsets.push_back(m->group_members.size());
  }
  m->member_offsets.assign(n + 1, 0);
  for (size_t cell : m->group_members) ++m->member_offsets[cell + 1];
  for (size_t k = 0; k < n; ++k) m->member_offsets[k + 1] += m->member_offsets[k];
  std::vector<size_t> cursor = m->member_offsets;
  m->member_groups.resize(m->group_members.size());
  m->group_volume.assign(n, 0);
  for (size_t g = 0; g < n; ++g)
    for (size_t q = m->group_offsets[g]; q < m->group_offsets[g + 1]; ++q) {
      const size_t k = m->group_members[q];
      m->member_groups[cursor[k]++] = g;
      m->group_volume[g] += m->cells[k].volume / (m->member_offsets[k + 1] - m->member_offsets[k]);
    }
  return true;
}

// Merge every cell below half its rectangle into same-level neighbours, smallest volume ratio
// first, always through the longest shared face (union-find with path halving).
bool merge_small_cells(StaticAmrMesh2D* m, std::string* error) {
  const size_t n = m->cells.size();
  std::vector<size_t> parent(n), order(n), root_id(n, n), count;
  std::iota(parent.begin(), parent.end(), 0);
  std::iota(order.begin(), order.end(), 0);
  std::vector<double> volume(n), target(n);
  for (size_t k = 0; k < n; ++k) {
    const AmrCell2D& c = m->cells[k];
    volume[k] = c.volume;
    target[k] = .5 * (c.x1 - c.x0) * (c.y1 - c.y0);
  }
  const auto root = [&](size_t k) {
    while (parent[k] != k) {
      parent[k] = parent[parent[k]];
      k = parent[k];
    }
    return k;
  };
  // std::sort (not stable_sort): tie order of equal ratios is part of the mesh contract.
  std::sort(order.begin(), order.end(),
            [&](size_t a, size_t b) { return volume[a] / target[a] < volume[b] / target[b]; });
  for (size_t seed : order) {
    size_t r = root(seed);
    while (volume[r] < target[r]) {
      size_t other = n;
      double best = 0;
      for (const AmrFace2D& f : m->faces) {
        if (f.left < 0 || f.right < 0 || m->cells[f.left].level != m->cells[f.right].level) continue;
        const size_t a = root(static_cast<size_t>(f.left)), b = root(static_cast<size_t>(f.right));
        if (a == b) continue;
        const size_t neighbor = a == r ? b : b == r ? a : n;
        if (neighbor < n && f.length > best) {
          other = neighbor;
          best = f.length;
        }
      }
      if (other == n) return fail(error, "small-cell agglomeration has no connected same-level neighbor");
      parent[r] = other;
      volume[other] += volume[r];
      target[other] = std::max(target[other], target[r]);
      r = other;
    }
  }
  std::vector<AmrCell2D> cells;
  std::vector<size_t> id(n);
  std::vector<long double> mx, my;
  for (size_t k = 0; k < n; ++k) {
    const size_t r = root(k);
    if (root_id[r] == n) {
      root_id[r] = cells.size();
      AmrCell2D c = m->cells[k];
      c.volume = 0;
      c.wall_begin = c.wall_count = 0;
      cells.push_back(c);
      count.push_back(0);
      mx.push_back(0);
      my.push_back(0);
    }
    const size_t g = root_id[r];
    id[k] = g;
    AmrCell2D& c = cells[g];
    const AmrCell2D& old = m->cells[k];
    c.x0 = std::min(c.x0, old.x0);
    c.x1 = std::max(c.x1, old.x1);
    c.y0 = std::min(c.y0, old.y0);
    c.y1 = std::max(c.y1, old.y1);
    c.ix = std::min(c.ix, old.ix);
    c.iy = std::min(c.iy, old.iy);
    c.volume += old.volume;
    mx[g] += static_cast<long double>(old.volume) * old.cx;
    my[g] += static_cast<long double>(old.volume) * old.cy;
    ++count[g];
    m->agglomeration.push_back({static_cast<double>(k), static_cast<double>(g), old.x0, old.x1, old.y0,
                                old.y1, old.volume, old.cx, old.cy});
  }
  for (size_t g = 0; g < cells.size(); ++g) {
    AmrCell2D& c = cells[g];
    if (count[g] > 1) {
      c.cx = static_cast<double>(mx[g] / c.volume);
      c.cy = static_cast<double>(my[g] / c.volume);
    }
  }
  std::vector<std::vector<WallFragment2D>> walls(cells.size());
  for (WallFragment2D w : m->walls) {
    w.owner_cell = id[w.owner_cell];
    walls[w.owner_cell].push_back(w);
  }
  m->walls.clear();
  for (size_t g = 0; g < cells.size(); ++g) {
    cells[g].wall_begin = m->walls.size();
    cells[g].wall_count = walls[g].size();
    m->walls.insert(m->walls.end(), walls[g].begin(), walls[g].end());
  }
  std::vector<AmrFace2D> faces;
  for (AmrFace2D f : m->faces) {
    if (f.left >= 0) f.left = static_cast<int>(id[f.left]);
    if (f.right >= 0) f.right = static_cast<int>(id[f.right]);
    if (f.left >= 0 && f.left == f.right) continue;  // now interior to a merged cell
    faces.push_back(f);
  }
  for (auto& pair : m->leaf_components) pair[1] = id[pair[1]];
  for (auto& leaf : m->partition)
    if (leaf[4] >= 0) leaf[4] = static_cast<int>(id[leaf[4]]);
  m->cells.swap(cells);
  m->faces.swap(faces);
  return true;
}

}  // namespace

int build_mesh(const StaticAmrOptions2D& o, const std::vector<double>& px, const std::vector<double>& py,
               const std::vector<int>& patches, StaticAmrMesh2D* m, std::string* error) {
  const auto invalid = [&](const std::string& message) {
    fail(error, message);
    return 1;
  };
  if (!m || o.nx < 1 || o.ny < 1 || o.levels < 0 || o.levels > 16 || o.buffer < 2 || o.max_cells < 1 ||
      o.max_cells > INT_MAX / 8 || !std::isfinite(o.xmin) || !std::isfinite(o.xmax) || !std::isfinite(o.ymin) ||
      !std::isfinite(o.ymax) || !(o.xmax > o.xmin) || !(o.ymax > o.ymin) ||
      static_cast<long long>(o.nx) * (1 << o.levels) > INT_MAX / 2 ||
      static_cast<long long>(o.ny) * (1 << o.levels) > INT_MAX / 2 || px.size() != py.size() ||
      (!px.empty() && px.size() < 3) || px.size() > INT_MAX || (!patches.empty() && patches.size() != px.size()))
    return invalid("invalid mesh options or polygon arrays");
  for (const AmrBox2D& box : o.boxes)
    if (box.level < 0 || box.level > o.levels || !std::isfinite(box.x0) || !std::isfinite(box.x1) ||
        !std::isfinite(box.y0) || !std::isfinite(box.y1) || !(box.x1 > box.x0) || !(box.y1 > box.y0))
      return invalid("invalid refinement box");
  const int scale = 1 << o.levels, npx = static_cast<int>(px.size());
  const double dx = (o.xmax - o.xmin) / (o.nx * scale), dy = (o.ymax - o.ymin) / (o.ny * scale);
  for (size_t a = 0; a < px.size(); ++a) {
    const size_t b = (a + 1) % px.size();
    if (!std::isfinite(px[a]) || !std::isfinite(py[a]) || (px[a] == px[b] && py[a] == py[b]) ||
        (!patches.empty() && patches[a] < 0))
      return invalid("invalid polygon vertex or patch");
    if (px[a] <= o.xmin + 3 * dx || px[a] >= o.xmax - 3 * dx || py[a] <= o.ymin + 3 * dy || py[a] >= o.ymax - 3 * dy)
      return invalid("body needs three finest cells of domain clearance");
  }
  *m = StaticAmrMesh2D{};
  m->options = o;
  m->body_x = px;
  m->body_y = py;
  const Polygon body{px.data(), py.data(), npx, patches.empty() ? nullptr : patches.data()};

  // 1. Refine each background cell by wall distance and boxes.
  std::vector<AmrCell2D> leaves;
  Refiner refiner{o, body, leaves};
  for (int j = 0; j < o.ny && !refiner.too_many; ++j)
    for (int i = 0; i < o.nx && !refiner.too_many; ++i) {
      AmrCell2D c = {};
      c.ix = i * scale;
      c.iy = j * scale;
      c.span = scale;
      refiner.visit(c);
    }
  if (refiner.too_many) return invalid("max-cells exceeded during refinement");

  // 2. Split leaves until neighbours differ by at most one level.
  std::vector<Link> links;
  for (;;) {
    links = pair_sides(leaves);
    std::vector<bool> split(leaves.size(), false);
    size_t count = 0;
    for (const Link& f : links)
      if (f.left >= 0 && f.right >= 0) {
        const int difference = leaves[f.left].level - leaves[f.right].level;
        if (std::abs(difference) > 1) {
          const int k = difference < 0 ? f.left : f.right;
          if (!split[k]) {
            split[k] = true;
            ++count;
          }
        }
      }
    if (count == 0) break;
    if (count > (o.max_cells - leaves.size()) / 3) return invalid("max-cells exceeded during 2:1 balance");
    std::vector<AmrCell2D> next;
    next.reserve(leaves.size() + 3 * count);
    for (size_t k = 0; k < leaves.size(); ++k)
      if (split[k]) split_leaf(leaves[k], o, &next);
      else next.push_back(leaves[k]);
    leaves.swap(next);
  }

  // 3. Cut every leaf by the body; each connected fluid region becomes a control volume.
  m->background_leaves = leaves.size();
  std::vector<int> first_cell(leaves.size(), -1);
  std::vector<std::vector<int>> regions_of(leaves.size());
  std::vector<std::vector<Point>> outlines;
  for (size_t k = 0; k < leaves.size(); ++k) {
    AmrCell2D c = leaves[k];
    std::vector<CutRegion> regions;
    const int status = rect_regions(body, c.x0, c.x1, c.y0, c.y1, &regions);
    if (status != 0) {
      std::ostringstream message;
      message.precision(17);
      message << "cut geometry status=" << status << " at leaf (" << c.ix << "," << c.iy << ") level=" << c.level
              << " bounds=[" << c.x0 << "," << c.x1 << "," << c.y0 << "," << c.y1 << "]";
      return invalid(message.str());
    }
    if (regions.size() > 1) ++m->split_leaves;
    for (CutRegion& r : regions) {
      if (r.volume == 0) continue;
      if (!r.walls.empty() && c.level != o.levels) return invalid("cut cell is not on finest level");
      c.volume = r.volume;
      c.cx = r.cx;
      c.cy = r.cy;
      c.wall_begin = m->walls.size();
      c.wall_count = r.walls.size();
      if (first_cell[k] < 0) first_cell[k] = static_cast<int>(m->cells.size());
      regions_of[k].push_back(static_cast<int>(m->cells.size()));
      m->leaf_components.push_back({k, m->cells.size()});
      for (WallFragment2D w : r.walls) {
        w.owner_cell = m->cells.size();
        m->walls.push_back(w);
      }
      m->cells.push_back(c);
      outlines.push_back(std::move(r.boundary));
    }
  }
  if (m->cells.empty()) return invalid("no active fluid cells");
  m->partition.reserve(leaves.size());
  for (size_t k = 0; k < leaves.size(); ++k) {
    const AmrCell2D& c = leaves[k];
    m->partition.push_back({c.ix, c.iy, c.span, c.level, first_cell[k]});
  }

  // 4. Faces: one per open subface, split further where a leaf holds several regions.
  for (const Link& link : links) {
    const int left = link.left < 0 ? -1 : first_cell[link.left];
    const int right = link.right < 0 ? -1 : first_cell[link.right];
    if (left < 0 && right < 0) continue;
    if ((link.left >= 0 && left < 0) || (link.right >= 0 && right < 0)) continue;  // solid neighbour
    const double origin = link.axis == 0 ? o.xmin : o.ymin, h = link.axis == 0 ? dx : dy;
    const double tangent_origin = link.axis == 0 ? o.ymin : o.xmin, tangent_h = link.axis == 0 ? dy : dx;
    AmrFace2D face = {};
    face.left = left;
    face.right = right;
    face.axis = link.axis;
    face.coordinate = origin + link.coordinate * h;
    face.lo = tangent_origin + link.lo * tangent_h;
    face.hi = tangent_origin + link.hi * tangent_h;
    face.boundary = link.left < 0 ? 2 * link.axis : link.right < 0 ? 2 * link.axis + 1 : -1;
    if (npx > 0) {
      open_aperture(body, link.axis, face.coordinate, face.lo, face.hi, &face.length, &face.centroid);
    } else {
      face.length = face.hi - face.lo;
      face.centroid = 0.5 * (face.lo + face.hi);
    }
    if (face.length == 0) continue;
    const auto has_outline = [&](int leaf) {
      return leaf >= 0 && !regions_of[leaf].empty() && !outlines[regions_of[leaf][0]].empty();
    };
    if (has_outline(link.left) || has_outline(link.right)) {
      // Assign each open interval to the region whose outline passes through its midpoint.
      const auto locate = [&](int leaf, double midpoint) {
        if (leaf < 0) return -1;
        const std::vector<int>& ids = regions_of[leaf];
        if (ids.size() == 1) return ids[0];
        const double x = link.axis == 0 ? face.coordinate : midpoint;
        const double y = link.axis == 1 ? face.coordinate : midpoint;
        const AmrCell2D& cell = leaves[leaf];
        const double tol = 1e-10 * std::min(cell.x1 - cell.x0, cell.y1 - cell.y0);
        for (int id : ids) {
          const std::vector<Point>& loop = outlines[id];
          for (size_t i = 0; i < loop.size(); ++i) {
            const Point a = loop[i], b = loop[(i + 1) % loop.size()];
            const double ex = b[0] - a[0], ey = b[1] - a[1], e2 = ex * ex + ey * ey;
            const double t = std::max(0., std::min(1., ((x - a[0]) * ex + (y - a[1]) * ey) / e2));
            if (std::hypot(x - a[0] - t * ex, y - a[1] - t * ey) <= tol) return id;
          }
        }
        return -2;
      };
      for (const Point iv : open_intervals(body, link.axis, face.coordinate, face.lo, face.hi)) {
        AmrFace2D sub = face;
        sub.lo = iv[0];
        sub.hi = iv[1];
        sub.length = sub.hi - sub.lo;
        sub.centroid = .5 * (sub.lo + sub.hi);
        sub.left = locate(link.left, sub.centroid);
        sub.right = locate(link.right, sub.centroid);
        if (sub.left == -2 || sub.right == -2)
          return invalid("cannot associate open aperture with a connected fluid component");
        if (sub.left >= 0 && sub.right >= 0 && m->cells[sub.left].level != m->cells[sub.right].level)
          return invalid("coarse/fine face touches a complex cut cell; increase buffer");
        m->faces.push_back(sub);
      }
      continue;
    }
    if (left >= 0 && right >= 0 && m->cells[left].level != m->cells[right].level) {
      if (m->cells[left].wall_count || m->cells[right].wall_count)
        return invalid("coarse/fine face touches cut cell; increase buffer");
      ++m->coarse_fine_faces;
    }
    m->faces.push_back(face);
  }

  // 5. Optional agglomeration, then CSR, closure audit and SRD groups.
  if (o.agglomerate_small && !merge_small_cells(m, error)) return 1;
  build_cell_faces(m);
  const size_t n = m->cells.size();
  for (size_t k = 0; k < n; ++k) {
    const AmrCell2D& c = m->cells[k];
    double closure[2] = {};
    for (size_t q = m->face_offsets[k]; q < m->face_offsets[k + 1]; ++q) {
      const AmrFace2D& f = m->faces[m->cell_faces[q]];
      closure[f.axis] += (f.left == static_cast<int>(k) ? 1 : -1) * f.length;
    }
    for (size_t q = c.wall_begin; q < c.wall_begin + c.wall_count; ++q) {
      closure[0] += m->walls[q].nx * m->walls[q].length;
      closure[1] += m->walls[q].ny * m->walls[q].length;
    }
    if (std::hypot(closure[0], closure[1]) > 1e-9 * std::max(c.x1 - c.x0, c.y1 - c.y0))
      return invalid("cell geometry is not closed");
  }
  return build_groups(m, error) ? 0 : 1;
}

int initialize_state(const StaticAmrMesh2D& m, const Params2D& p, const std::array<double, 4>& w,
                     StaticAmrState2D* s) {
  if (!s || !(p.gamma > 1) || !std::isfinite(p.gamma) || !std::isfinite(w[0]) || !std::isfinite(w[1]) ||
      !std::isfinite(w[2]) || !std::isfinite(w[3]) || !(w[0] > 0) || !(w[3] > 0))
    return 1;
  *s = StaticAmrState2D{};
  const size_t n = m.cells.size();
  s->u.resize((m.transport.sa_model ? 5 : 4) * n);
  const double values[4] = {w[0], w[0] * w[1], w[0] * w[2], w[3] / (p.gamma - 1) + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2])};
  for (int v = 0; v < 4; ++v) std::fill(s->u.begin() + v * n, s->u.begin() + (v + 1) * n, values[v]);
  if (m.transport.sa_model) std::fill(s->u.begin() + 4 * n, s->u.end(), w[0] * m.transport.sa_nu_inf);
  return 0;
}

}  // namespace hipamr
}  // namespace cfd
