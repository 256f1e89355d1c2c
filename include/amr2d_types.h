// amr2d_types.h -- data types of the static cut-cell AMR time step, layout-identical to the
// reference's so that a caller built against the reference headers can link this library.
//
// Every struct below mirrors one reference type field for field (same order, same widths):
//   Params2D, Stretch ........ src/euler2d.h:33-52
//   WallFragment2D ........... src/euler2d.h:87-91
//   AmrBoundary2D ............ src/static_amr2d_boundary.h:14-21
//   AmrTransport2D, AmrGradientTerm2D, AmrWallSample2D, AmrViscousFaceGeometry2D,
//   AmrViscousTable2D ........ src/static_amr2d_viscous.h:15-37
//   AmrBox2D, StaticAmrOptions2D, AmrCell2D, AmrFace2D, StaticAmrMesh2D, StaticAmrState2D
//                              src/static_amr2d.h:29-71
// The layout is part of the drop-in contract: static_amr_gpu_create() receives these objects
// by reference from code compiled against the reference's own headers.  Only plain data lives
// in namespace cfd; all functions of this project live in cfd::hipamr to keep clear of the
// reference's inline helpers (no ODR overlap when both are linked into one executable).
#ifndef HIPAMR_AMR2D_TYPES_H
#define HIPAMR_AMR2D_TYPES_H

#include <array>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace cfd {

// Option enumerations (src/types.h:7-13, src/euler2d.h:23-24).
enum FluxType { FLUX_HLLC = 0, FLUX_AUSMPUP = 1 };
enum LimiterType { LIM_NONE = 0, LIM_MINMOD = 1, LIM_VENKAT = 2, LIM_FIRST_ORDER = 3 };
enum BCType { BC_PERIODIC = 0, BC_TRANSMISSIVE = 1, BC_REFLECT = 2 };
enum TimeIntegrator2D { TIME_SSP_RK3 = 0, TIME_BACKWARD_EULER = 1 };

struct Stretch {
  int type;
  double param;
};

// Solver parameters.  Only gamma, cfl, flux, limiter, ausm_mco, bc, dt_fixed and
// time_integrator are read by the time step; the rest keep the reference layout.
struct Params2D {
  double gamma;
  double cfl;
  int flux;
  int limiter;
  double venkat_k;
  double ausm_mco;
  int bc[4];
  double dt_fixed;
  int time_integrator;
  double pseudo_cfl;
  int implicit_maxiter;
  double implicit_rtol, implicit_atol;
  Stretch sx, sy;
};

// A straight piece of the body boundary inside one cut cell; (nx, ny) points fluid -> solid.
struct WallFragment2D {
  size_t owner_cell;
  int source_edge_id, body_id, patch_id;
  double x0, y0, x1, y1, nx, ny, length;
};

// Boundary kinds of the four domain sides (src/static_amr2d_boundary.h:12-13).
enum AmrBoundaryKind {
  AMR_BC_INHERIT = -1, AMR_BC_TRANSMISSIVE = 1, AMR_BC_REFLECT = 2, AMR_BC_FARFIELD = 3,
  AMR_BC_TOTAL_INLET = 4, AMR_BC_PRESSURE_OUTLET = 5, AMR_BC_NOSLIP = 6
};
struct AmrBoundary2D {
  int kind = AMR_BC_INHERIT;
  double reference[4] = {1, 1.8, 0, 1.0 / 1.4};  // rho, u, v, p (also the outlet backflow state)
  double total_pressure = 0, total_temperature = 0, gas_constant = 1;
  double direction[2] = {1, 0};
  double outlet_pressure = 0;
  double wall_velocity[2] = {0, 0};
};

enum AmrFlowModel { AMR_MODEL_EULER = 0, AMR_MODEL_LAMINAR = 1 };
enum AmrViscosityLaw { AMR_MU_CONSTANT = 0, AMR_MU_SUTHERLAND = 1 };
enum AmrViscousStabilization { AMR_VISCOUS_VOLUME = 0, AMR_VISCOUS_SRD = 1 };
struct AmrTransport2D {
  int model = AMR_MODEL_EULER, law = AMR_MU_CONSTANT;
  int stabilization = AMR_VISCOUS_VOLUME;
  double mu_ref = 0, gas_constant = 1, temperature_ref = 1, sutherland = 0, viscous_cfl = .1;
  int sa_model = 0;  // 0 off, 1 SA-neg
  double sa_nu_inf = 0, sa_cfl = .1;
  double wall_probe_min_h = 0;
};
// Least-squares gradient term: source >= 0 neighbour cell, -1 body wall, -2 no-slip domain side.
struct AmrGradientTerm2D { int source, side; double x, y; };
struct AmrWallSample2D { int donor; double distance; };
struct AmrViscousFaceGeometry2D { double dx, dy, ex, ey, distance; };
struct AmrViscousTable2D {
  std::vector<size_t> offsets;
  std::vector<AmrGradientTerm2D> terms;
  std::vector<AmrWallSample2D> wall_samples, face_samples;
  std::vector<AmrViscousFaceGeometry2D> face_geometry;
  std::vector<double> wall_distance;
};

struct AmrBox2D { double x0, x1, y0, y1; int level; };
struct StaticAmrOptions2D {
  int nx = 16, ny = 12, levels = 3, buffer = 4;
  double xmin = -4, xmax = 4, ymin = -3, ymax = 3;
  size_t max_cells = 2000000;
  bool agglomerate_small = false;
  std::vector<AmrBox2D> boxes;
};
struct AmrCell2D {
  int ix, iy, span, level;
  double x0, x1, y0, y1, volume, cx, cy;
  size_t wall_begin, wall_count;
};
// One shared subface: left/right are cell ids (-1 outside), the normal points along +axis,
// boundary is the domain side 0..3 (xmin, xmax, ymin, ymax) or -1 for an interior face.
struct AmrFace2D {
  int left, right, axis, boundary;
  double coordinate, lo, hi, length, centroid;
};
struct StaticAmrMesh2D {
  StaticAmrOptions2D options;
  AmrBoundary2D boundaries[4];
  AmrTransport2D transport;
  AmrViscousTable2D viscous;
  std::vector<double> body_x, body_y;
  std::vector<AmrCell2D> cells;
  std::vector<std::array<int, 5>> partition;
  std::vector<std::array<size_t, 2>> leaf_components;
  size_t split_leaves = 0;
  std::vector<std::array<double, 9>> agglomeration;
  std::vector<AmrFace2D> faces;
  std::vector<WallFragment2D> walls;
  std::vector<size_t> face_offsets, cell_faces;
  std::vector<size_t> group_offsets, group_members, member_offsets, member_groups;
  std::vector<double> group_volume;
  size_t background_leaves = 0, coarse_fine_faces = 0, small_cells = 0;
};
// Conserved state, structure of arrays: u[var * cells + cell], var = rho, rho*u, rho*v, rho*E
// and, with SA-neg, rho*nutilde.
struct StaticAmrState2D {
  std::vector<double> u;
  double time = 0;
  long steps = 0, rejected = 0;
};

static_assert(sizeof(Params2D) == 136, "Params2D layout");
static_assert(sizeof(WallFragment2D) == 80, "WallFragment2D layout");
static_assert(sizeof(AmrBoundary2D) == 104, "AmrBoundary2D layout");
static_assert(sizeof(AmrTransport2D) == 88, "AmrTransport2D layout");
static_assert(sizeof(AmrGradientTerm2D) == 24, "AmrGradientTerm2D layout");
static_assert(sizeof(AmrWallSample2D) == 16, "AmrWallSample2D layout");
static_assert(sizeof(AmrViscousFaceGeometry2D) == 40, "AmrViscousFaceGeometry2D layout");
static_assert(sizeof(AmrCell2D) == 88, "AmrCell2D layout");
static_assert(sizeof(AmrFace2D) == 56, "AmrFace2D layout");

}  // namespace cfd

#endif  // HIPAMR_AMR2D_TYPES_H
