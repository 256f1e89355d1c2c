// ugrid_axis.cpp -- geometry of one axis of the tensor-product grid (ugrid_axis in ugrid_hip.h).
//
// Face coordinates from the stretching law (reference grid2d_stretch_faces,
// src/euler2d.cpp:38-62), then cell centres and widths with two ghost layers per side
// (build_axis, src/euler2d.cpp:66-105): a uniform axis uses the exact width (b - a) / n and the
// linear centre formula everywhere, ghosts included; a stretched axis takes centres and widths
// from the faces and translates (periodic) or mirrors (otherwise) them into the ghosts.  The
// steps and their order are the reference's, so the geometry is bit-identical to it.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ugrid_hip.h"

namespace {

constexpr int kGhosts = 2;
enum { STRETCH_UNIFORM = 0, STRETCH_GEOMETRIC = 1, STRETCH_TANH_ENDS = 2, STRETCH_SINH_CENTER = 3 };
constexpr int BC_PERIODIC = 0;

int report(char* err, size_t len, const std::string& message) {
  if (err && len) {
    const size_t k = message.size() < len - 1 ? message.size() : len - 1;
    std::memcpy(err, message.data(), k);
    err[k] = 0;
  }
  return 1;
}

void faces(int n, double a, double b, int type, double param, double* xf) {
  const double length = b - a;
  for (int i = 0; i <= n; ++i) {
    const double xi = static_cast<double>(i) / n;
    double f = xi;
    if (type == STRETCH_GEOMETRIC) {
      if (std::fabs(param - 1.0) > 1e-12) f = (std::pow(param, i) - 1.0) / (std::pow(param, n) - 1.0);
    } else if (type == STRETCH_TANH_ENDS) {
      if (param > 1e-12) f = 0.5 * (1.0 + std::tanh(param * (2.0 * xi - 1.0)) / std::tanh(param));
    } else if (type == STRETCH_SINH_CENTER) {
      if (param > 1e-12) f = 0.5 * (1.0 + std::sinh(param * (2.0 * xi - 1.0)) / std::sinh(param));
    }
    xf[i] = a + length * f;
  }
  xf[0] = a;
  xf[n] = b;
}

}  // namespace

extern "C" int ugrid_axis(int32_t n, double a, double b, int32_t stretch_type, double stretch_param, int32_t bc_lo,
                          int32_t bc_hi, double* xf, double* xc, double* dx, char* err, size_t err_len) {
  if (n < 1 || !xf || !xc || !dx || !std::isfinite(a) || !std::isfinite(b) || !(b > a) || stretch_type < 0 ||
      stretch_type > 3 || !std::isfinite(stretch_param))
    return report(err, err_len, "invalid grid axis");
  if ((bc_lo == BC_PERIODIC) != (bc_hi == BC_PERIODIC))
    return report(err, err_len, "periodic boundary must be set on both sides of an axis");
  const int ng = kGhosts;
  faces(n, a, b, stretch_type, stretch_param, xf);
  if (stretch_type == STRETCH_UNIFORM) {
    const double h = (b - a) / n;
    for (int i = -ng; i < n + ng; ++i) {
      xc[i + ng] = a + (i + 0.5) * h;
      dx[i + ng] = h;
    }
    return 0;
  }
  for (int i = 0; i < n; ++i) {
    xc[i + ng] = 0.5 * (xf[i] + xf[i + 1]);
    dx[i + ng] = xf[i + 1] - xf[i];
  }
  const double length = b - a;
  for (int k = 1; k <= ng; ++k) {
    if (bc_lo == BC_PERIODIC) {
      xc[ng - k] = xc[ng + n - k] - length;
      dx[ng - k] = dx[ng + n - k];
      xc[ng + n - 1 + k] = xc[ng + k - 1] + length;
      dx[ng + n - 1 + k] = dx[ng + k - 1];
    } else {
      xc[ng - k] = 2.0 * a - xc[ng + k - 1];
      dx[ng - k] = dx[ng + k - 1];
      xc[ng + n - 1 + k] = 2.0 * b - xc[ng + n - k];
      dx[ng + n - 1 + k] = dx[ng + n - k];
    }
  }
  return 0;
}
