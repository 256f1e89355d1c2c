// ugrid_caller.cpp -- a caller of the reference's legacy-grid GPU backend API, compiled against
// include/euler2d_gpu.h and linked against libamr_hip.so (built by build()).  It sets up a
// Solver2D the way solver2d_init does (src/euler2d.cpp:107-151; geometry from ugrid_axis), fills
// one Riemann problem per case along x, and calls cfd::solver2d_run_cuda.
//
//   ugrid_caller layout                  print the 50 layout values of this header's Solver2D / Cut2D
//   ugrid_caller probe                   print solver2d_cuda_available's answer
//   ugrid_caller run OUT nx ny ncase flux bc0 bc1 bc2 bc3 stretch param tend maxsteps [limiter [R [be]]]
//       write OUT: int64 status, nsteps, nreject; then the initial U, the final U, t, dt (doubles).
//       R > 0: a solid disc of radius R (48 vertices) at (0.47, 0.53) as cut cells (Cut2D filled
//       from ugrid_cut_build); be = 1: backward Euler.
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "euler2d_gpu.h"
#include "ugrid_hip.h"

using namespace cfd;

static void layout() {
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
  for (int64_t x : v) std::printf("%lld\n", static_cast<long long>(x));
}

static bool axis(int n, double a, double b, int type, double param, int lo, int hi, Array& xf, Array& xc,
                 Array& dx) {
  xf.allocate(n + 1);
  xc.allocate(n + 4);
  dx.allocate(n + 4);
  char err[256];
  if (ugrid_axis(n, a, b, type, param, lo, hi, xf.data(), xc.data(), dx.data(), err, sizeof(err))) {
    std::fprintf(stderr, "axis: %s\n", err);
    return false;
  }
  return true;
}

int main(int argc, char** argv) {
  if (argc >= 2 && !std::strcmp(argv[1], "layout")) {
    layout();
    return 0;
  }
  if (argc >= 2 && !std::strcmp(argv[1], "probe")) {
    char reason[256];
    const bool ok = solver2d_cuda_available(reason, sizeof(reason));
    std::printf("%d %s\n", ok ? 1 : 0, reason);
    return 0;
  }
  if (argc < 15 || std::strcmp(argv[1], "run")) {
    std::fprintf(stderr, "usage: ugrid_caller layout | probe | run OUT nx ny ncase flux bc0 bc1 bc2 bc3 stretch param "
                         "tend maxsteps [limiter [R [be]]]\n");
    return 2;
  }
  const char* out = argv[2];
  const int nx = std::atoi(argv[3]), ny = std::atoi(argv[4]), ncase = std::atoi(argv[5]);
  Solver2D s{};
  Params2D& p = s.par;
  p.gamma = 1.4;
  p.cfl = 0.4;
  p.flux = std::atoi(argv[6]);
  p.limiter = argc > 15 ? std::atoi(argv[15]) : LIM_FIRST_ORDER;
  p.venkat_k = 5.0;
  p.ausm_mco = 1.0;
  for (int k = 0; k < 4; ++k) p.bc[k] = std::atoi(argv[7 + k]);
  p.dt_fixed = 0.0;
  p.time_integrator = 0;
  p.pseudo_cfl = 0.4;
  p.implicit_maxiter = 2000;
  p.implicit_rtol = 1e-8;
  p.implicit_atol = 1e-11;
  p.sx = {std::atoi(argv[11]), std::atof(argv[12])};
  p.sy = {0, 0.0};
  const double tend = std::atof(argv[13]);
  const long maxsteps = std::atol(argv[14]);
  const double radius = argc > 16 ? std::atof(argv[16]) : 0.0;
  if (argc > 17 && std::atoi(argv[17]) == 1) {
    p.time_integrator = 1;
    p.cfl = 2.0;
  }

  Grid2D& g = s.grid;
  g.nx = nx;
  g.ny = ny;
  g.ng = 2;
  g.mx = nx + 4;
  g.my = ny + 4;
  g.xmin = 0.0;
  g.xmax = 1.0;
  g.ymin = 0.0;
  g.ymax = 1.0;
  if (!axis(nx, 0.0, 1.0, p.sx.type, p.sx.param, p.bc[0], p.bc[1], g.xf, g.xc, g.dx) ||
      !axis(ny, 0.0, 1.0, p.sy.type, p.sy.param, p.bc[2], p.bc[3], g.yf, g.yc, g.dy))
    return 2;
  s.ncase = ncase;
  const size_t states = static_cast<size_t>(ncase) * NVAR2D * g.mx * g.my;
  for (Array* a : {&s.U, &s.U1, &s.U2, &s.R, &s.W, &s.SX, &s.SY}) a->allocate(states);
  s.FX.allocate(static_cast<size_t>(ncase) * NVAR2D * (nx + 1) * ny);
  s.FY.allocate(static_cast<size_t>(ncase) * NVAR2D * nx * (ny + 1));
  s.t.allocate(ncase);
  s.dt.allocate(ncase);
  s.mco.allocate(ncase);
  s.ib = nullptr;
  s.cut = nullptr;
  Cut2D cut;
  if (radius > 0.0) {
    std::vector<double> px(48), py(48);
    for (int k = 0; k < 48; ++k) {
      const double a = 0.0131 + 2.0 * 3.141592653589793 * k / 48;
      px[k] = 0.47 + radius * std::cos(a);
      py[k] = 0.53 + radius * std::sin(a);
    }
    ugrid_axes axes{nx, ny, g.xf.data(), g.xc.data(), g.dx.data(), g.yf.data(), g.yc.data(), g.dy.data()};
    ugrid_cut* built = nullptr;
    char err[256];
    if (ugrid_cut_build(&axes, px.data(), py.data(), 48, nullptr, &built, err, sizeof(err))) {
      std::fprintf(stderr, "cut: %s\n", err);
      return 2;
    }
    ugrid_cut_view v{};
    ugrid_cut_get(built, &v);
    const size_t cells = static_cast<size_t>(g.mx) * g.my, nfx = static_cast<size_t>(nx + 1) * ny,
                 nfy = static_cast<size_t>(nx) * (ny + 1), ns = static_cast<size_t>(v.nsmall);
    auto fill = [](Array& a, const double* src, size_t n) {
      a.allocate(n);
      for (size_t k = 0; k < n; ++k) a[k] = src[k];
    };
    cut.nsmall = v.nsmall;
    cut.nbad_nbhd = v.nbad_nbhd;
    cut.nsolid = v.nsolid;
    cut.ncut = v.ncut;
    fill(cut.vol, v.vol, cells);
    fill(cut.cx, v.cx, cells);
    fill(cut.cy, v.cy, cells);
    fill(cut.flag, v.flag, cells);
    fill(cut.axl, v.axl, nfx);
    fill(cut.axy, v.axy, nfx);
    fill(cut.ayl, v.ayl, nfy);
    fill(cut.ayx, v.ayx, nfy);
    fill(cut.wax, v.wax, cells);
    fill(cut.way, v.way, cells);
    fill(cut.wcx, v.wcx, cells);
    fill(cut.wcy, v.wcy, cells);
    fill(cut.nbrcount, v.nbrcount, cells);
    fill(cut.smallpos, v.smallpos, cells);
    fill(cut.sidx, v.sidx, ns);
    fill(cut.scnt, v.scnt, ns);
    fill(cut.smem, v.smem, 25 * ns);
    fill(cut.svhat, v.svhat, ns);
    fill(cut.sxhat, v.sxhat, ns);
    fill(cut.syhat, v.syhat, ns);
    ugrid_cut_destroy(built);
    s.cut = &cut;
  }
  // One Riemann problem per case (left state for x < 0.5), every case its own reference Mach.
  const double table[4][8] = {{1.0, 0.0, 0.0, 1.0, 0.125, 0.0, 0.0, 0.1},
                              {1.0, 0.75, 0.1, 1.0, 0.125, 0.0, -0.1, 0.1},
                              {1.0, -2.0, 0.0, 0.4, 1.0, 2.0, 0.0, 0.4},
                              {1.0, 0.0, 0.3, 1000.0, 1.0, 0.0, -0.2, 0.01}};
  for (int c = 0; c < ncase; ++c) {
    s.mco[c] = 0.2 + 0.2 * c;
    const double* st = table[c % 4];
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double* w = g.xc[i + 2] < 0.5 ? st : st + 4;
        const size_t cells = static_cast<size_t>(g.mx) * g.my;
        const size_t k = static_cast<size_t>(c) * NVAR2D * cells + static_cast<size_t>(j + 2) * g.mx + (i + 2);
        s.U[k] = w[0];
        s.U[k + cells] = w[0] * w[1];
        s.U[k + 2 * cells] = w[0] * w[2];
        s.U[k + 3 * cells] = w[3] / (p.gamma - 1.0) + 0.5 * w[0] * w[1] * w[1] + 0.5 * w[0] * w[2] * w[2];
      }
  }
  std::vector<double> initial(s.U.data(), s.U.data() + states);
  const int status = solver2d_run_cuda(&s, tend, maxsteps);
  FILE* f = std::fopen(out, "wb");
  if (!f) return 2;
  const int64_t head[3] = {status, s.nsteps, s.nreject};
  std::fwrite(head, sizeof(head), 1, f);
  std::fwrite(initial.data(), sizeof(double), states, f);
  std::fwrite(s.U.data(), sizeof(double), states, f);
  std::fwrite(s.t.data(), sizeof(double), ncase, f);
  std::fwrite(s.dt.data(), sizeof(double), ncase, f);
  std::fclose(f);
  std::printf("status %d nsteps %ld nreject %ld\n", status, s.nsteps, s.nreject);
  return 0;
}
