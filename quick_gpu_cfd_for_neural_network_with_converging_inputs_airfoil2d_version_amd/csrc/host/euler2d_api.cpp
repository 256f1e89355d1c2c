// euler2d_api.cpp -- the reference's legacy GPU backend entry points (include/euler2d_gpu.h)
// over the C ABI of ugrid_hip.h:
//   cfd::solver2d_run_cuda        replaces src/euler2d_cuda.cu:530-708
//   cfd::solver2d_cuda_available  replaces src/euler2d_cuda.cu:510-528
//
// The caller's Solver2D comes from code compiled against the reference's src/euler2d.h (or this
// project's layout-identical include/euler2d_gpu.h).  This file names the type only (the symbol
// names depend on nothing else) and reads it through a private view of the same layout, so no
// definition of the reference's classes or their inline functions lives in the library.
#include <chrono>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "amr2d_types.h"
#include "ugrid_hip.h"

namespace cfd {
struct Solver2D;
int solver2d_run_cuda(Solver2D* s, double tend, long maxsteps);
bool solver2d_cuda_available(char* reason, size_t reason_size);
}  // namespace cfd

namespace {

// Views of src/array.h:55-57, src/euler2d.h:54-60, 118-135.
struct ArrayView {
  double* p;
  size_t n;
};
struct GridView {
  int nx, ny, ng, mx, my;
  double xmin, xmax, ymin, ymax;
  ArrayView xf, yf, xc, dx, yc, dy;
};
struct ImplicitView {
  long updates, retries;
  double initial_defect, final_defect, tolerance;
};
struct VectorView {  // std::vector (libstdc++): begin, end, capacity end
  void* p[3];
};
struct CutView {  // src/euler2d.h:98-117
  VectorView wall_fragments, wall_cells;
  int nsmall, nbad_nbhd, nsolid, ncut;
  ArrayView vol, cx, cy, flag, axl, axy, ayl, ayx, wax, way, wcx, wcy, nbrcount, smallpos, sidx, scnt, smem, svhat,
      sxhat, syhat;
};
struct SolverView {
  GridView grid;
  cfd::Params2D par;
  int ncase;
  ArrayView U, U1, U2, R, W, SX, SY, FX, FY, t, dt, mco;
  long nsteps, nreject;
  ImplicitView implicit;
  const void* ib;
  const void* cut;
};
static_assert(sizeof(ArrayView) == 16 && sizeof(GridView) == 152 && sizeof(SolverView) == 560, "Solver2D layout");
static_assert(sizeof(CutView) == 384, "Cut2D layout");

}  // namespace

bool cfd::solver2d_cuda_available(char* reason, size_t reason_size) {
  char text[256];
  const bool ok = ugrid_device_available(text, sizeof(text)) != 0;
  if (reason != nullptr && reason_size > 0) std::snprintf(reason, reason_size, "%s", text);
  return ok;
}

int cfd::solver2d_run_cuda(Solver2D* solver, double tend, long maxsteps) {
  SolverView& s = *reinterpret_cast<SolverView*>(solver);
  const bool implicit = s.par.time_integrator == TIME_BACKWARD_EULER;
  if (!implicit && s.par.time_integrator != TIME_SSP_RK3) return 3;
  if (implicit && (s.par.limiter != LIM_FIRST_ORDER || s.ib != nullptr)) {
    std::fprintf(stderr, "backward Euler currently requires first-order regular/cut-cell grids (no legacy IB)\n");
    return 3;
  }
  char available[256];
  if (!solver2d_cuda_available(available, sizeof(available))) {
    std::fprintf(stderr, "CUDA backend unavailable: %s\n", available);
    return 3;
  }
  if (s.par.limiter != LIM_FIRST_ORDER) {
    std::fprintf(stderr,
                 "CUDA backend currently requires --limiter first; refusing to change the method implicitly\n");
    return 3;
  }
  if (s.ib != nullptr) {
    std::fprintf(stderr, "CUDA backend supports regular grids and cut-cell+SRD, not the legacy IB path\n");
    return 3;
  }
  const CutView* cut = static_cast<const CutView*>(s.cut);
  if (cut != nullptr && cut->nbad_nbhd != 0) {
    std::fprintf(stderr, "CUDA backend refuses %d invalid SRD neighborhoods\n", cut->nbad_nbhd);
    return 3;
  }
  const GridView& g = s.grid;
  const size_t states = static_cast<size_t>(s.ncase) * 4 * g.mx * g.my;
  if (g.ng != 2 || g.mx != g.nx + 4 || g.my != g.ny + 4 || s.U.n != states || g.dx.n != static_cast<size_t>(g.mx) ||
      g.dy.n != static_cast<size_t>(g.my) || s.t.n != static_cast<size_t>(s.ncase) ||
      s.dt.n != static_cast<size_t>(s.ncase) || s.mco.n != static_cast<size_t>(s.ncase)) {
    std::fprintf(stderr, "HIP backend: solver arrays do not match the grid (solver2d_init not called?)\n");
    return 3;
  }
  ugrid_setup setup{};
  setup.nx = g.nx;
  setup.ny = g.ny;
  setup.ncase = s.ncase;
  setup.dx = g.dx.p;
  setup.dy = g.dy.p;
  ugrid_params& p = setup.params;
  p.gamma = s.par.gamma;
  p.cfl = s.par.cfl;
  p.flux = s.par.flux;
  p.limiter = s.par.limiter;
  p.ausm_mco = s.par.ausm_mco;
  for (int k = 0; k < 4; ++k) p.bc[k] = s.par.bc[k];
  p.dt_fixed = s.par.dt_fixed;
  p.time_integrator = s.par.time_integrator;
  p.implicit_maxiter = s.par.implicit_maxiter;
  p.pseudo_cfl = s.par.pseudo_cfl;
  p.implicit_rtol = s.par.implicit_rtol;
  p.implicit_atol = s.par.implicit_atol;
  setup.mco = s.mco.p;
  ugrid_cut_view cv{};
  if (cut != nullptr) {
    const size_t cells = static_cast<size_t>(g.mx) * g.my;
    const size_t nfx = static_cast<size_t>(g.nx + 1) * g.ny, nfy = static_cast<size_t>(g.nx) * (g.ny + 1);
    const size_t ns = static_cast<size_t>(cut->nsmall);
    if (cut->vol.n != cells || cut->flag.n != cells || cut->wax.n != cells || cut->way.n != cells ||
        cut->nbrcount.n != cells || cut->smallpos.n != cells || cut->axl.n != nfx || cut->ayl.n != nfy ||
        cut->scnt.n != ns || cut->svhat.n != ns || cut->smem.n != 25 * ns) {
      std::fprintf(stderr, "HIP backend: cut-cell arrays do not match the grid\n");
      return 3;
    }
    cv.nsmall = cut->nsmall;
    cv.nbad_nbhd = cut->nbad_nbhd;
    cv.nsolid = cut->nsolid;
    cv.ncut = cut->ncut;
    cv.vol = cut->vol.p;
    cv.cx = cut->cx.p;
    cv.cy = cut->cy.p;
    cv.flag = cut->flag.p;
    cv.axl = cut->axl.p;
    cv.axy = cut->axy.p;
    cv.ayl = cut->ayl.p;
    cv.ayx = cut->ayx.p;
    cv.wax = cut->wax.p;
    cv.way = cut->way.p;
    cv.wcx = cut->wcx.p;
    cv.wcy = cut->wcy.p;
    cv.nbrcount = cut->nbrcount.p;
    cv.smallpos = cut->smallpos.p;
    cv.sidx = cut->sidx.p;
    cv.scnt = cut->scnt.p;
    cv.smem = cut->smem.p;
    cv.svhat = cut->svhat.p;
    cv.sxhat = cut->sxhat.p;
    cv.syhat = cut->syhat.p;
    setup.cut = &cv;
  }

  char err[512] = {0};
  const auto t0 = std::chrono::steady_clock::now();
  ugrid_session* session = nullptr;
  if (ugrid_create(&setup, s.U.p, states, s.t.p, s.dt.p, &session, err, sizeof(err))) {
    std::fprintf(stderr, "HIP backend: %s\n", err);
    return 3;
  }
  const auto t1 = std::chrono::steady_clock::now();
  ugrid_meta meta{};
  int status = ugrid_run(session, tend, maxsteps, &meta, err, sizeof(err));
  const auto t2 = std::chrono::steady_clock::now();
  ugrid_implicit_info info{};
  const bool have_info = ugrid_implicit_stats(session, &info) == 0;
  if (status == 2 && !implicit) std::fprintf(stderr, "CUDA solver: %s\n", err);
  if ((status == 2 || status == 4) && implicit) std::fprintf(stderr, "CUDA %s\n", err);
  // State, clocks and steps come back after every device run (a backward-Euler step that
  // cannot advance included); refused controls stop before the device (src/euler2d_cuda.cu:694-707).
  const bool dt_stop = implicit && status == 3 && have_info && info.reason == 5;
  if (status == 0 || status == 1 || status == 2 || status == 4 || dt_stop) {
    char derr[256];
    if (ugrid_download(session, s.U.p, states, s.t.p, s.dt.p, derr, sizeof(derr))) {
      std::fprintf(stderr, "HIP backend: %s\n", derr);
      status = 3;
    } else {
      s.nsteps += static_cast<long>(meta.nsteps);
      s.nreject += static_cast<long>(meta.nreject);
      if (have_info) {
        s.implicit.updates += static_cast<long>(info.updates);
        s.implicit.retries += static_cast<long>(info.retries);
        if (info.checks > 0) {
          s.implicit.initial_defect = info.initial_defect;
          s.implicit.final_defect = info.final_defect;
          s.implicit.tolerance = info.tolerance;
        }
      }
    }
  } else {
    std::fprintf(stderr, "HIP backend: %s\n", err);
    status = 3;
  }
  ugrid_destroy(session);
  std::fprintf(stderr, "hip: device=%s setup_h2d=%.6f s solve=%.6f s\n", available,
               std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count());
  return status;
}

// Offsets of the Solver2D and Cut2D fields this library reads, for the layout check of the tests:
// sizes of Solver2D, Grid2D, Array, Params2D, ImplicitStats2D, then the offsets of grid,
// grid.{nx, ng, mx, my, xmin, xf, yf, xc, dx, yc, dy}, par, ncase, U, U1, U2, R, W, SX, SY, FX,
// FY, t, dt, mco, nsteps, nreject, implicit, ib, cut (36 values), then sizeof(Cut2D) and the
// offsets of nsmall, nbad_nbhd, vol, flag, axl, ayl, wax, way, nbrcount, smallpos, scnt, smem,
// svhat (14 values).
extern "C" int ugrid_solver2d_layout(int64_t* out, int32_t capacity) {
  const int64_t v[] = {
      sizeof(SolverView), sizeof(GridView), sizeof(ArrayView), sizeof(cfd::Params2D), sizeof(ImplicitView),
      offsetof(SolverView, grid), offsetof(GridView, nx), offsetof(GridView, ng), offsetof(GridView, mx),
      offsetof(GridView, my), offsetof(GridView, xmin), offsetof(GridView, xf), offsetof(GridView, yf),
      offsetof(GridView, xc), offsetof(GridView, dx), offsetof(GridView, yc), offsetof(GridView, dy),
      offsetof(SolverView, par), offsetof(SolverView, ncase), offsetof(SolverView, U), offsetof(SolverView, U1),
      offsetof(SolverView, U2), offsetof(SolverView, R), offsetof(SolverView, W), offsetof(SolverView, SX),
      offsetof(SolverView, SY), offsetof(SolverView, FX), offsetof(SolverView, FY), offsetof(SolverView, t),
      offsetof(SolverView, dt), offsetof(SolverView, mco), offsetof(SolverView, nsteps),
      offsetof(SolverView, nreject), offsetof(SolverView, implicit), offsetof(SolverView, ib),
      offsetof(SolverView, cut), sizeof(CutView), offsetof(CutView, nsmall), offsetof(CutView, nbad_nbhd),
      offsetof(CutView, vol), offsetof(CutView, flag), offsetof(CutView, axl), offsetof(CutView, ayl),
      offsetof(CutView, wax), offsetof(CutView, way), offsetof(CutView, nbrcount), offsetof(CutView, smallpos),
      offsetof(CutView, scnt), offsetof(CutView, smem), offsetof(CutView, svhat)};
  const int32_t n = static_cast<int32_t>(sizeof(v) / sizeof(v[0]));
  for (int32_t k = 0; k < n && k < capacity; ++k) out[k] = v[k];
  return n;
}
