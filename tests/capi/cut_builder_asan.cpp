// cut_builder_asan.cpp -- the native cut-cell builder (ugrid_cut_build, csrc/host/ugrid_cut.cpp)
// under AddressSanitizer and UndefinedBehaviorSanitizer on the host: every view array is read end
// to end on several grids and bodies, and the refusals (a polygon too small, an edge on a grid
// line, a body too close to a side) return their codes without touching memory they do not own.
// Built and run by tests/test_capi_euler2d.py with g++ (host code only).
#include <cmath>
#include <cstdio>
#include <vector>

#include "ugrid_hip.h"

namespace {

struct Axes {
  std::vector<double> xf, xc, dx, yf, yc, dy;
  ugrid_axes view{};
};

bool make_axes(int nx, int ny, int sx, double px, Axes* a) {
  a->xf.resize(nx + 1);
  a->xc.resize(nx + 4);
  a->dx.resize(nx + 4);
  a->yf.resize(ny + 1);
  a->yc.resize(ny + 4);
  a->dy.resize(ny + 4);
  char err[256];
  if (ugrid_axis(nx, -1.0, 1.0, sx, px, 1, 1, a->xf.data(), a->xc.data(), a->dx.data(), err, sizeof(err)) ||
      ugrid_axis(ny, -0.8, 0.8, 0, 0.0, 2, 2, a->yf.data(), a->yc.data(), a->dy.data(), err, sizeof(err))) {
    std::printf("axis: %s\n", err);
    return false;
  }
  a->view = ugrid_axes{nx, ny, a->xf.data(), a->xc.data(), a->dx.data(), a->yf.data(), a->yc.data(), a->dy.data()};
  return true;
}

double touch(const double* p, size_t n) {
  double s = 0.0;
  for (size_t k = 0; k < n; ++k) s += std::fabs(p[k]) < 1e300 ? p[k] : 0.0;
  return s;
}

int build(const Axes& a, const std::vector<double>& px, const std::vector<double>& py, int expect) {
  ugrid_cut* cut = nullptr;
  char err[256] = {0};
  const int status = ugrid_cut_build(&a.view, px.data(), py.data(), static_cast<int32_t>(px.size()), nullptr, &cut, err,
                                     sizeof(err));
  if (status != expect) {
    std::printf("status %d, expected %d: %s\n", status, expect, err);
    return 1;
  }
  if (status) {
    if (cut) return 1;  // a refusal hands back no geometry
    return 0;
  }
  ugrid_cut_view v{};
  if (ugrid_cut_get(cut, &v)) return 1;
  const size_t cells = static_cast<size_t>(a.view.nx + 4) * (a.view.ny + 4);
  const size_t nfx = static_cast<size_t>(a.view.nx + 1) * a.view.ny, nfy = static_cast<size_t>(a.view.nx) * (a.view.ny + 1);
  const size_t ns = static_cast<size_t>(v.nsmall);
  double s = 0.0;
  for (const double* p : {v.vol, v.cx, v.cy, v.flag, v.wax, v.way, v.wcx, v.wcy, v.nbrcount, v.smallpos}) s += touch(p, cells);
  for (const double* p : {v.axl, v.axy}) s += touch(p, nfx);
  for (const double* p : {v.ayl, v.ayx}) s += touch(p, nfy);
  for (const double* p : {v.sidx, v.scnt, v.svhat, v.sxhat, v.syhat}) s += touch(p, ns);
  s += touch(v.smem, 25 * ns);
  for (uint64_t k = 0; k < v.n_walls; ++k) s += v.walls[k].length;
  ugrid_cut_destroy(cut);
  if (!(s == s) || v.ncut <= 0 || v.nsolid <= 0) return 1;
  std::printf("ok nx=%d ny=%d n=%zu cut=%d solid=%d small=%d walls=%llu\n", a.view.nx, a.view.ny, px.size(), v.ncut,
              v.nsolid, v.nsmall, static_cast<unsigned long long>(v.n_walls));
  return 0;
}

void ellipse(double cx, double cy, double rx, double ry, int n, std::vector<double>* px, std::vector<double>* py) {
  px->resize(n);
  py->resize(n);
  for (int k = 0; k < n; ++k) {
    const double t = 0.0137 + 2.0 * 3.141592653589793 * k / n;
    (*px)[k] = cx + rx * std::cos(t);
    (*py)[k] = cy + ry * std::sin(t);
  }
}

}  // namespace

int main() {
  int bad = 0;
  std::vector<double> px, py;
  for (const auto& g : {std::vector<double>{48, 40, 0, 0.0}, std::vector<double>{37, 52, 1, 1.02},
                        std::vector<double>{131, 97, 2, 1.2}}) {
    Axes a;
    if (!make_axes(static_cast<int>(g[0]), static_cast<int>(g[1]), static_cast<int>(g[2]), g[3], &a)) return 2;
    ellipse(0.1, 0.05, 0.33, 0.27, 64, &px, &py);
    bad += build(a, px, py, 0);
    ellipse(-0.07, 0.0, 0.52, 0.11, 96, &px, &py);
    bad += build(a, px, py, 0);
    bad += build(a, {-0.45, 0.357, 0.357}, {0.013, -0.313, 0.341}, 0);
    bad += build(a, {0.0, 0.1}, {0.0, 0.1}, 1);          // fewer than three vertices
    ellipse(0.0, 0.0, 0.98, 0.5, 64, &px, &py);          // reaches the side columns
    bad += build(a, px, py, 2);
  }
  Axes a;
  if (!make_axes(40, 40, 0, 0.0, &a)) return 2;
  bad += build(a, {-0.5, 0.5, 0.5}, {0.0, -0.3, 0.3}, 3);  // x = 0.5 is a grid line of this axis
  std::printf(bad ? "FAILED %d\n" : "all ok\n", bad);
  return bad ? 1 : 0;
}
