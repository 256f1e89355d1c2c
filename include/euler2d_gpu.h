// euler2d_gpu.h -- the reference's legacy tensor-product-grid solver types and its GPU backend
// entry points, implemented for MI355X (gfx950) by libamr_hip.so.
//
// Layout-identical to the reference (checked against it by the tests), so a caller compiled
// against the reference's src/euler2d.h links this library unchanged, and a caller without the
// reference can compile against this header:
//   Array .................... src/array.h:12-58
//   Grid2D ................... src/euler2d.h:54-60
//   ImplicitStats2D .......... src/euler2d.h:118-121
//   WallCell2D, Cut2D ........ src/euler2d.h:92-94, 98-117
//   Solver2D ................. src/euler2d.h:123-135 (IB2D only by pointer)
//   Params2D, Stretch, enums . amr2d_types.h (src/euler2d.h:23-52, src/types.h:7-13)
//   solver2d_run_cuda ........ replaces src/euler2d.h:211 (src/euler2d_cuda.cu:530-708)
//   solver2d_cuda_available .. replaces src/euler2d.h:213 (src/euler2d_cuda.cu:510-528)
// solver2d_run_cuda returns 0 (every case at tend), 1 (maxsteps steps), 2 (nonphysical
// candidate after 12 retries), 3 (unavailable, unsupported or invalid), 4 (backward-Euler step
// not converged); it reads the grid, parameters, U, t, dt and mco, and writes back U (ghost cells
// as the reference's device array ends), t, dt, nsteps, nreject and the implicit statistics.
// Supported as by the reference's backend: FP64, first-order limiter, SSP-RK3 and backward
// Euler on regular and cut-cell grids (SRD after each stage); the IB path returns 3.
#ifndef HIPAMR_EULER2D_GPU_H
#define HIPAMR_EULER2D_GPU_H

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "amr2d_types.h"

namespace cfd {

enum { NVAR2D = 4 };
enum SolverBackend { BACKEND_CPU = 0, BACKEND_CUDA = 1 };
enum StretchType { STRETCH_UNIFORM = 0, STRETCH_GEOMETRIC = 1, STRETCH_TANH_ENDS = 2, STRETCH_SINH_CENTER = 3 };

// A zero-initialised double array that owns its memory (same members and behaviour as the
// reference's: calloc on allocate, abort on failure, move-only).
class Array {
 public:
  Array() : p_(nullptr), n_(0) {}
  explicit Array(size_t n) : p_(nullptr), n_(0) { allocate(n); }
  ~Array() { release(); }
  Array(const Array&) = delete;
  Array& operator=(const Array&) = delete;
  Array(Array&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  Array& operator=(Array&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void allocate(size_t n) {
    release();
    p_ = static_cast<double*>(std::calloc(n > 0 ? n : 1, sizeof(double)));
    if (p_ == nullptr) {
      std::fprintf(stderr, "Array::allocate: out of memory (%zu doubles)\n", n);
      std::abort();
    }
    n_ = n;
  }
  void release() {
    std::free(p_);
    p_ = nullptr;
    n_ = 0;
  }
  double* data() { return p_; }
  const double* data() const { return p_; }
  size_t size() const { return n_; }
  double& operator[](size_t i) { return p_[i]; }
  const double& operator[](size_t i) const { return p_[i]; }

 private:
  double* p_;
  size_t n_;
};

struct Grid2D {
  int nx, ny, ng, mx, my;
  double xmin, xmax, ymin, ymax;
  Array xf, yf;  // face coordinates, nx + 1 and ny + 1
  Array xc, dx;  // cell centres and widths with ghosts, mx
  Array yc, dy;  // my
};

struct IB2D;

struct WallCell2D {
  size_t owner_cell, begin, count;
};

// Cut-cell geometry and SRD neighbourhoods (built by the reference's cut2d_build_polygon, or by
// ugrid_cut_build of ugrid_hip.h).  Cell arrays have mx * my entries, x-face arrays
// (nx + 1) * ny (index j * (nx + 1) + i), y-face arrays nx * (ny + 1) (index j * nx + i).
struct Cut2D {
  std::vector<WallFragment2D> wall_fragments;
  std::vector<WallCell2D> wall_cells;
  int nsmall;
  int nbad_nbhd;
  int nsolid, ncut;
  Array vol;
  Array cx, cy;
  Array flag;
  Array axl, axy;
  Array ayl, ayx;
  Array wax, way;
  Array wcx, wcy;
  Array nbrcount;
  Array smallpos;
  Array sidx;
  Array scnt;
  Array smem;
  Array svhat, sxhat, syhat;
};

struct ImplicitStats2D {
  long updates, retries;
  double initial_defect, final_defect, tolerance;
};

struct Solver2D {
  Grid2D grid;
  Params2D par;
  int ncase;
  Array U, U1, U2;  // conserved states U[((c * 4 + v) * my + j + ng) * mx + i + ng]
  Array R;
  Array W;
  Array SX, SY;
  Array FX, FY;
  Array t, dt, mco;  // per case: time, last step, AUSM+-up reference Mach number
  long nsteps;
  long nreject;
  ImplicitStats2D implicit;
  const IB2D* ib;
  const Cut2D* cut;
};

int solver2d_run_cuda(Solver2D* s, double tend, long maxsteps);
bool solver2d_cuda_available(char* reason, size_t reason_size);

}  // namespace cfd

#endif
