// cutcell.h -- polygon / axis-aligned rectangle geometry for the cut-cell mesher.
//
// Follows the reference's geometry primitives (src/cutcell2d.cpp:27-245, src/sdf.h:15-42)
// operation for operation: the mesh arrays must come out bit-identical, so every rounding
// step (division order, hypot/atan2, long double moments) is kept.
#ifndef HIPAMR_CUTCELL_H
#define HIPAMR_CUTCELL_H

#include <array>
#include <vector>

#include "amr2d_types.h"

namespace cfd {
namespace hipamr {

using Point = std::array<double, 2>;

// The body polygon: closed polyline, vertex k -> k+1 (last -> first), optional patch per edge.
struct Polygon {
  const double* x;
  const double* y;
  int n;
  const int* patch;  // may be null
};

// Cut result of one rectangular leaf (reference RectMetrics2D, src/static_amr2d.h:12-16).
struct CutRegion {
  double volume = 0, cx = 0, cy = 0;
  std::vector<WallFragment2D> walls;
  std::vector<Point> boundary;  // traced outline, only for leaves split into several regions
};

double segment_distance2(double x, double y, double ax, double ay, double bx, double by);
bool inside_polygon(const Polygon& body, double x, double y);

// Fluid aperture of the segment {axis-coordinate = c, lo < t < hi}; length 0 when closed.
void open_aperture(const Polygon& body, int axis, double c, double lo, double hi, double* length,
                   double* centroid);
// The same without snapping near-full / near-empty lengths (src/cutcell2d.cpp:49-77).
void open_aperture_raw(const Polygon& body, int axis, double c, double lo, double hi, double* length,
                       double* centroid);
// Liang-Barsky clip of the segment to the rectangle (src/cutcell2d.cpp:79-103); false when
// nothing of parameter length above 1e-14 remains.
bool clip_to_rect(double* ax, double* ay, double* bx, double* by, double x0, double x1, double y0,
                  double y1);
// Unit normal of edge a -> a + 1 pointing into the fluid, decided at (mx, my) + eps * normal
// (src/cutcell2d.cpp:105-115).
void fluid_normal(const Polygon& body, int a, double mx, double my, double eps, double* nx, double* ny);
// Fluid sub-intervals of the same segment.
std::vector<Point> open_intervals(const Polygon& body, int axis, double c, double lo, double hi);
// Distance from the rectangle to the polygon boundary (0 when an edge enters it).
double rect_wall_distance(const Polygon& body, double x0, double x1, double y0, double y1);
// Single-region cut metrics; status 0 ok, 2 several wall arcs, 3 inconsistent area.
int rect_cut(const Polygon& body, double x0, double x1, double y0, double y1, CutRegion* out);
// All fluid regions of a leaf; status 0 ok, 2..5 unsupported topology.
int rect_regions(const Polygon& body, double x0, double x1, double y0, double y1,
                 std::vector<CutRegion>* out);

}  // namespace hipamr
}  // namespace cfd

#endif
