// amr_oracle.cpp -- TEST INFRASTRUCTURE ONLY.  FP64 CPU restatement of the reference's
// explicit cut-cell time step, used by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker of the HIP path.  Nothing in the product links or calls it.
//
// Each function follows the reference CPU code operation for operation (the oracle is pinned
// bit for bit against the compiled reference in the build container, oracle/_ref):
//   primitive ........................ src/static_amr2d.cpp:143-154
//   residual (faces, gather, coupling) src/static_amr2d.cpp:156-206
//   redistribute (SRD) ............... src/static_amr2d.cpp:208-226
//   run (CFL dt, SSP-RK3, rejection) . src/static_amr2d.cpp:384-432
//   rhs .............................. src/static_amr2d.cpp:369-377
//   viscous evaluate ................. src/static_amr2d_viscous.cpp:77-142
//   SA-neg evaluate .................. src/static_amr2d_sa.cpp:17-45
//   HLLC / AUSM+-up / exact flux ..... src/flux.h:19-176
//   boundary states .................. src/static_amr2d_boundary.h:22-102
//   viscosity, stress, gradients ..... src/static_amr2d_viscous.h:141-179
//   SA-neg point model ............... src/sa_model.h:12-39
//   steady monitor (residual norms,
//   wall loads, CL/CD/CM) ............ src/quick_airfoil.cpp:75-85
// The C ABI (oracle_*) takes the mesh as an amr_mesh_view (include/amr_hip.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "amr_hip.h"

namespace oracle {

enum { BC_TRANSMISSIVE = 1, BC_REFLECT = 2, BC_FARFIELD = 3, BC_TOTAL_INLET = 4, BC_PRESSURE_OUTLET = 5, BC_NOSLIP = 6 };
enum { MODEL_EULER = 0, MODEL_LAMINAR = 1, MU_CONSTANT = 0, MU_SUTHERLAND = 1, VISCOUS_VOLUME = 0, VISCOUS_SRD = 1 };
enum { FLUX_HLLC = 0, FLUX_AUSMPUP = 1 };

using Vec = std::vector<double>;

// ------------------------------------------------------------------ fluxes (src/flux.h) --
void exact_flux(double rho, double un, double ut, double p, double gamma, double F[4]) {
  const double E = p / (gamma - 1.0) + 0.5 * rho * un * un + 0.5 * rho * ut * ut;
  F[0] = rho * un;
  F[1] = rho * un * un + p;
  F[2] = rho * un * ut;
  F[3] = un * (E + p);
}

// HLLC with Einfeldt/Roe wave speeds; the tangential velocity is advected.
void hllc(double rL, double uL, double vL, double pL, double rR, double uR, double vR, double pR, double g,
          double F[4]) {
  const double aL = std::sqrt(g * pL / rL), aR = std::sqrt(g * pR / rR);
  const double EL = pL / (g - 1.0) + 0.5 * rL * uL * uL + 0.5 * rL * vL * vL;
  const double ER = pR / (g - 1.0) + 0.5 * rR * uR * uR + 0.5 * rR * vR * vR;
  const double sL = std::sqrt(rL), sR = std::sqrt(rR);
  const double HL = (EL + pL) / rL, HR = (ER + pR) / rR;
  const double uT = (sL * uL + sR * uR) / (sL + sR);
  const double vT = (sL * vL + sR * vR) / (sL + sR);
  const double HT = (sL * HL + sR * HR) / (sL + sR);
  const double aT = std::sqrt(std::max((g - 1.0) * (HT - 0.5 * uT * uT - 0.5 * vT * vT), 1e-300));
  const double SL = std::min(uL - aL, uT - aT);
  const double SR = std::max(uR + aR, uT + aT);
  double FL[4], FR[4];
  exact_flux(rL, uL, vL, pL, g, FL);
  exact_flux(rR, uR, vR, pR, g, FR);
  if (SL >= 0.0) {
    for (int k = 0; k < 4; ++k) F[k] = FL[k];
    return;
  }
  if (SR <= 0.0) {
    for (int k = 0; k < 4; ++k) F[k] = FR[k];
    return;
  }
  const double Ss = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR));
  if (Ss >= 0.0) {
    const double c = rL * (SL - uL) / (SL - Ss);
    const double U3 = c * (EL / rL + (Ss - uL) * (Ss + pL / (rL * (SL - uL))));
    F[0] = FL[0] + SL * (c - rL);
    F[1] = FL[1] + SL * (c * Ss - rL * uL);
    F[2] = FL[2] + SL * (c * vL - rL * vL);
    F[3] = FL[3] + SL * (U3 - EL);
  } else {
    const double c = rR * (SR - uR) / (SR - Ss);
    const double U3 = c * (ER / rR + (Ss - uR) * (Ss + pR / (rR * (SR - uR))));
    F[0] = FR[0] + SR * (c - rR);
    F[1] = FR[1] + SR * (c * Ss - rR * uR);
    F[2] = FR[2] + SR * (c * vR - rR * vR);
    F[3] = FR[3] + SR * (U3 - ER);
  }
}

// AUSM+-up (Liou 2006) with Kp = 0.25, Ku = 0.75, sigma = 1, beta = 1/8.
double mach_1(double M, int s) { return 0.5 * (M + s * std::fabs(M)); }
double mach_2(double M, int s) { return s * 0.25 * (M + s) * (M + s); }
double mach_4(double M, int s, double beta) {
  if (std::fabs(M) >= 1.0) return mach_1(M, s);
  return mach_2(M, s) * (1.0 - s * 16.0 * beta * mach_2(M, -s));
}
double pressure_5(double M, int s, double alpha) {
  if (std::fabs(M) >= 1.0) return 0.5 * (1.0 + s * (M > 0.0 ? 1.0 : (M < 0.0 ? -1.0 : 0.0)));
  return mach_2(M, s) * ((s * 2.0 - M) - s * 16.0 * alpha * M * mach_2(M, -s));
}
void ausm(double rL, double uL, double vL, double pL, double rR, double uR, double vR, double pR, double g,
          double Mco, double F[4]) {
  const double Kp = 0.25, Ku = 0.75, sigma = 1.0, beta = 0.125;
  const double aL = std::sqrt(g * pL / rL), aR = std::sqrt(g * pR / rR);
  const double a12 = 0.5 * (aL + aR), rho12 = 0.5 * (rL + rR);
  const double ML = uL / a12, MR = uR / a12;
  const double Mbar2 = 0.5 * (uL * uL + uR * uR) / (a12 * a12);
  const double M02 = std::min(1.0, std::max(Mbar2, Mco * Mco));
  const double M0 = std::sqrt(M02);
  const double fa = M0 * (2.0 - M0);
  const double alpha = 3.0 / 16.0 * (-4.0 + 5.0 * fa * fa);
  const double M12 = mach_4(ML, +1, beta) + mach_4(MR, -1, beta) -
                     (Kp / fa) * std::max(1.0 - sigma * Mbar2, 0.0) * (pR - pL) / (rho12 * a12 * a12);
  const double Pp = pressure_5(ML, +1, alpha), Pm = pressure_5(MR, -1, alpha);
  const double p12 = Pp * pL + Pm * pR - Ku * Pp * Pm * (rL + rR) * (fa * a12) * (uR - uL);
  const bool up = M12 > 0.0;
  const double mdot = a12 * M12 * (up ? rL : rR);
  const double rho = up ? rL : rR, u = up ? uL : uR, v = up ? vL : vR, p = up ? pL : pR;
  const double E = p / (g - 1.0) + 0.5 * rho * u * u + 0.5 * rho * v * v;
  const double H = (E + p) / rho;
  F[0] = mdot;
  F[1] = mdot * u + p12;
  F[2] = mdot * v;
  F[3] = mdot * H;
}

// ------------------------------------------- boundary states (src/static_amr2d_boundary.h) --
bool primitive_valid(const double* w) {
  return std::isfinite(w[0]) && w[0] > 0 && std::isfinite(w[1]) && std::isfinite(w[2]) && std::isfinite(w[3]) && w[3] > 0;
}
int side_kind(const amr_mesh_view& m, const amr_params& p, int side) {
  return m.boundaries[side].kind < 0 ? p.bc[side] : m.boundaries[side].kind;
}
bool boundary_valid(const amr_boundary& b, int legacy, int side) {
  if (b.kind < -1) return false;
  const int kind = b.kind < 0 ? legacy : b.kind;
  if (kind == BC_TRANSMISSIVE || kind == BC_REFLECT) return true;
  if (kind == BC_NOSLIP)
    return std::isfinite(b.wall_velocity[0]) && std::isfinite(b.wall_velocity[1]) && b.wall_velocity[side / 2] == 0;
  if (kind < BC_FARFIELD || kind > BC_PRESSURE_OUTLET || !primitive_valid(b.reference)) return false;
  if (kind == BC_PRESSURE_OUTLET) return std::isfinite(b.outlet_pressure) && b.outlet_pressure > 0;
  if (kind == BC_TOTAL_INLET) {
    const double d2 = b.direction[0] * b.direction[0] + b.direction[1] * b.direction[1];
    const double dn = (side % 2 ? 1.0 : -1.0) * b.direction[side / 2];
    return std::isfinite(d2) && std::fabs(d2 - 1) < 1e-12 && dn < 0 && std::isfinite(b.total_pressure) &&
           b.total_pressure > 0 && std::isfinite(b.total_temperature) && b.total_temperature > 0 &&
           std::isfinite(b.gas_constant) && b.gas_constant > 0;
  }
  return true;
}
// Exterior Riemann state wb from the interior state wi (global velocity components).
bool boundary_state(const amr_boundary& b, int legacy, double gamma, int side, const double* wi, double* wb) {
  const int axis = side / 2, normal = axis + 1, tangent = 2 - axis;
  const double sign = side % 2 ? 1.0 : -1.0, gm = gamma - 1;
  const int kind = b.kind < 0 ? legacy : b.kind;
  for (int v = 0; v < 4; ++v) wb[v] = wi[v];
  if (kind == BC_TRANSMISSIVE) return true;
  if (kind == BC_NOSLIP) {
    wb[1] = 2 * b.wall_velocity[0] - wi[1];
    wb[2] = 2 * b.wall_velocity[1] - wi[2];
    return true;
  }
  if (kind == BC_REFLECT) {
    wb[normal] = -wi[normal];
    return true;
  }
  if (!primitive_valid(wi) || !(gamma > 1)) return false;
  const double ai = std::sqrt(gamma * wi[3] / wi[0]), un = sign * wi[normal];
  if (kind == BC_TOTAL_INLET) {
    if (!(un < 0 && un > -ai)) return false;
    const double dn = sign * b.direction[axis], A = gamma * b.gas_constant * b.total_temperature;
    if (!(dn < 0) || !std::isfinite(A)) return false;
    const double rp = un + 2 * ai / gm, upper = std::sqrt(2 * A / (gm + 2 * dn * dn));
    const double f0 = 2 * std::sqrt(A) / gm - rp;
    const double f1 = dn * upper + 2 * std::sqrt(A - .5 * gm * upper * upper) / gm - rp;
    if (!(f0 >= 0 && f1 <= 0)) return false;
    double lo = 0, hi = upper;
    for (int it = 0; it < 64; ++it) {
      const double v = .5 * (lo + hi), a = std::sqrt(A - .5 * gm * v * v);
      if (dn * v + 2 * a / gm > rp) lo = v;
      else hi = v;
    }
    const double v = .5 * (lo + hi), temp = (A - .5 * gm * v * v) / (gamma * b.gas_constant);
    const double pressure = b.total_pressure * std::pow(temp / b.total_temperature, gamma / gm);
    wb[0] = pressure / (b.gas_constant * temp);
    wb[1] = v * b.direction[0];
    wb[2] = v * b.direction[1];
    wb[3] = pressure;
    return primitive_valid(wb);
  }
  if (kind == BC_PRESSURE_OUTLET && un >= 0) {
    if (un >= ai) return true;
    const double density = wi[0] * std::pow(b.outlet_pressure / wi[3], 1 / gamma);
    const double ab = std::sqrt(gamma * b.outlet_pressure / density), ub = un + 2 * (ai - ab) / gm;
    if (ub >= 0) {
      wb[0] = density;
      wb[normal] = sign * ub;
      wb[3] = b.outlet_pressure;
      return primitive_valid(wb);
    }
  }
  if (kind != BC_FARFIELD && kind != BC_PRESSURE_OUTLET) return false;
  if (un >= ai) return true;
  if (un <= -ai) {
    for (int v = 0; v < 4; ++v) wb[v] = b.reference[v];
    return primitive_valid(wb);
  }
  const double* wr = b.reference;
  const double ar = std::sqrt(gamma * wr[3] / wr[0]);
  const double rp = un + 2 * ai / gm, rm = sign * wr[normal] - 2 * ar / gm;
  const double ab = .25 * gm * (rp - rm), ub = .5 * (rp + rm);
  if (!(ab > 0) || !std::isfinite(ab)) return false;
  const double advrho = un >= 0 ? wi[0] : wr[0], advp = un >= 0 ? wi[3] : wr[3];
  const double advt = un >= 0 ? wi[tangent] : wr[tangent];
  const double entropy = advp / std::pow(advrho, gamma);
  wb[0] = std::pow(ab * ab / (gamma * entropy), 1 / gm);
  wb[3] = entropy * std::pow(wb[0], gamma);
  wb[normal] = sign * ub;
  wb[tangent] = advt;
  return primitive_valid(wb);
}

// ----------------------------------------------- transport (src/static_amr2d_viscous.h) --
double viscosity(const amr_transport& t, double temperature) {
  if (!(temperature > 0) || !std::isfinite(temperature)) return -1;
  if (t.law == MU_CONSTANT) return t.mu_ref;
  const double ratio = temperature / t.temperature_ref;
  return t.mu_ref * ratio * std::sqrt(ratio) * (t.temperature_ref + t.sutherland) / (temperature + t.sutherland);
}
void viscous_flux(double mu, const double* g, double nx, double ny, double u, double v, double* f) {
  const double dil = (2.0 / 3) * (g[0] + g[3]);
  const double t0 = mu * (2 * g[0] - dil), t1 = mu * (g[1] + g[2]), t2 = mu * (2 * g[3] - dil);
  f[0] = 0;
  f[1] = t0 * nx + t1 * ny;
  f[2] = t1 * nx + t2 * ny;
  f[3] = u * f[1] + v * f[2];
}
void face_gradient(const double* L, const double* R, double dx, double dy, double du, double dv, double* g) {
  const double d2 = dx * dx + dy * dy;
  for (int q = 0; q < 4; ++q) g[q] = .5 * (L[q] + R[q]);
  const double cu = (du - g[0] * dx - g[1] * dy) / d2, cv = (dv - g[2] * dx - g[3] * dy) / d2;
  g[0] += cu * dx;
  g[1] += cu * dy;
  g[2] += cv * dx;
  g[3] += cv * dy;
}
void wall_gradient(double nx, double ny, double d, double du, double dv, double* g) {
  g[0] = nx * du / d;
  g[1] = ny * du / d;
  g[2] = nx * dv / d;
  g[3] = ny * dv / d;
}

// ------------------------------------------------------------ SA-neg (src/sa_model.h) --
constexpr double cb1 = .1355, cb2 = .622, sigma = 2.0 / 3, kappa = .41;
constexpr double cw1 = cb1 / (kappa * kappa) + (1 + cb2) / sigma;
struct SaPoint {
  double fv1, modified_vorticity, fw, source, source_rate, diffusion;
};
double sa_eddy(double working, double molecular) {
  if (working <= 0) return 0;
  const double chi = working / molecular, cube = chi * chi * chi;
  return working * cube / (cube + 7.1 * 7.1 * 7.1);
}
SaPoint sa_point(double working, double molecular, double distance, double omega) {
  SaPoint a{};
  const double chi = working / molecular, cube = chi * chi * chi, d2 = distance * distance;
  if (working < 0) {
    const double fn = (16 + cube) / (16 - cube);
    a.diffusion = (molecular + working * fn) / sigma;
    a.source = cb1 * (1 - 1.2) * omega * working + cw1 * working * working / d2;
    a.source_rate = cb1 * .2 * omega + 2 * cw1 * std::fabs(working) / d2;
    return a;
  }
  a.fv1 = cube / (cube + 7.1 * 7.1 * 7.1);
  const double fv2 = 1 - chi / (1 + chi * a.fv1);
  const double sbar = working * fv2 / (kappa * kappa * d2);
  a.modified_vorticity = sbar >= -.7 * omega ? omega + sbar : omega * ((-.01 * omega - .1 * sbar) / (-.5 * omega - sbar));
  const double r = a.modified_vorticity == 0 ? 10 : std::fmin(working / (a.modified_vorticity * kappa * kappa * d2), 10.);
  const double r6 = r * r * r * r * r * r, g = r + .3 * (r6 - r), g6 = g * g * g * g * g * g;
  a.fw = g * std::pow(65 / (g6 + 64), 1.0 / 6);
  const double ft2 = 1.2 * std::exp(-.5 * chi * chi);
  const double production = cb1 * (1 - ft2) * a.modified_vorticity, destruction = cw1 * a.fw - cb1 * ft2 / (kappa * kappa);
  a.source = production * working - destruction * working * working / d2;
  a.source_rate = std::fabs(production) + 2 * std::fabs(destruction) * working / d2;
  a.diffusion = (molecular + working) / sigma;
  return a;
}
double sa_source(const SaPoint& a, double rho, double gx, double gy, double rx, double ry) {
  return rho * (a.source + cb2 / sigma * (gx * gx + gy * gy)) - a.diffusion * (rx * gx + ry * gy);
}

// -------------------------------------------------------------- the time step operators --
struct ViscousFields {
  Vec gradient, mu, face_flux, wall_flux, residual, cell_dt;
};

inline size_t csr_count(const uint64_t* offsets, size_t k) { return offsets[k + 1] - offsets[k]; }

bool primitive(const Vec& u, size_t n, double gamma, Vec* w) {
  w->resize(4 * n);
  int bad = 0;
#pragma omp parallel for reduction(| : bad)
  for (size_t k = 0; k < n; ++k) {
    const double rho = u[k], mx = u[n + k], my = u[2 * n + k], energy = u[3 * n + k];
    if (!std::isfinite(rho) || !(rho > 0) || !std::isfinite(mx) || !std::isfinite(my) || !std::isfinite(energy)) {
      bad = 1;
      continue;
    }
    const double vx = mx / rho, vy = my / rho, p = (gamma - 1) * (energy - 0.5 * (mx * vx + my * vy));
    if (!std::isfinite(p) || !(p > 0)) {
      bad = 1;
      continue;
    }
    (*w)[k] = rho;
    (*w)[n + k] = vx;
    (*w)[2 * n + k] = vy;
    (*w)[3 * n + k] = p;
  }
  return bad == 0;
}

int viscous_evaluate(const amr_mesh_view& m, const amr_params& p, const Vec& w, ViscousFields* out, const Vec* u) {
  const amr_transport& t = m.transport;
  if (!out || t.model != MODEL_LAMINAR) return 1;
  const size_t n = m.n_cells, nf = m.n_faces, nw = m.n_walls;
  if (w.size() != 4 * n) return 1;
  ViscousFields& a = *out;
  a.gradient.assign(4 * n, 0);
  a.mu.resize(n);
  a.face_flux.assign(4 * nf, 0);
  a.wall_flux.assign(4 * nw, 0);
  a.residual.assign(4 * n, 0);
  a.cell_dt.resize(n);
  Vec fc(nf), wc(nw), eddy(n, 0);
  int bad = 0;
  if (t.sa_model && (!u || u->size() != 5 * n)) return 1;
#pragma omp parallel for reduction(| : bad)
  for (size_t k = 0; k < n; ++k) {
    a.mu[k] = viscosity(t, w[3 * n + k] / (w[k] * t.gas_constant));
    if (!(a.mu[k] > 0) || !std::isfinite(a.mu[k])) bad = 1;
    if (t.sa_model) {
      const double ph = (*u)[4 * n + k] / w[k];
      if (!std::isfinite(ph)) {
        bad = 1;
        continue;
      }
      eddy[k] = w[k] * sa_eddy(ph, a.mu[k] / w[k]);
    }
    for (size_t q = m.visc_offsets[k]; q < m.visc_offsets[k + 1]; ++q) {
      const amr_grad_term& term = m.visc_terms[q];
      double uu = 0, vv = 0;
      if (term.source >= 0) {
        uu = w[n + term.source];
        vv = w[2 * n + term.source];
      } else if (term.source == -2) {
        uu = m.boundaries[term.side].wall_velocity[0];
        vv = m.boundaries[term.side].wall_velocity[1];
      }
      const double du = uu - w[n + k], dv = vv - w[2 * n + k];
      a.gradient[k] += term.x * du;
      a.gradient[n + k] += term.y * du;
      a.gradient[2 * n + k] += term.x * dv;
      a.gradient[3 * n + k] += term.y * dv;
    }
  }
  if (bad) return 1;
#pragma omp parallel for reduction(| : bad)
  for (size_t q = 0; q < nf; ++q) {
    const amr_face& f = m.faces[q];
    const amr_face_geometry& geo = m.face_geometry[q];
    const int k = f.left >= 0 ? f.left : f.right;
    double g[4], mu = a.mu[k] + eddy[k], uf = w[n + k], vf = w[2 * n + k], nu = mu / w[k], distance = geo.distance;
    for (int v = 0; v < 4; ++v) g[v] = a.gradient[v * n + k];
    if (f.boundary < 0) {
      const int l = f.left, r = f.right;
      double gl[4], gr[4];
      for (int v = 0; v < 4; ++v) {
        gl[v] = a.gradient[v * n + l];
        gr[v] = a.gradient[v * n + r];
      }
      face_gradient(gl, gr, geo.dx, geo.dy, w[n + r] - w[n + l], w[2 * n + r] - w[2 * n + l], g);
      uf = .5 * (w[n + l] + w[n + r]) + g[0] * geo.ex + g[1] * geo.ey;
      vf = .5 * (w[2 * n + l] + w[2 * n + r]) + g[2] * geo.ex + g[3] * geo.ey;
      const double temp = .5 * (w[3 * n + l] / w[l] + w[3 * n + r] / w[r]) / t.gas_constant;
      mu = viscosity(t, temp) + .5 * (eddy[l] + eddy[r]);
      nu = std::max((a.mu[l] + eddy[l]) / w[l], (a.mu[r] + eddy[r]) / w[r]);
    } else {
      distance = m.face_samples[q].distance;
      const int bc = side_kind(m, p, f.boundary);
      const double sign = f.boundary % 2 ? 1 : -1;
      if (bc == BC_NOSLIP) {
        mu = a.mu[k];
        uf = m.boundaries[f.boundary].wall_velocity[0];
        vf = m.boundaries[f.boundary].wall_velocity[1];
        wall_gradient(f.axis == 0 ? sign : 0, f.axis == 1 ? sign : 0, distance, uf - w[n + k], vf - w[2 * n + k], g);
      } else if (bc == BC_REFLECT) {
        g[1] = 0;
        g[2] = 0;
        if (f.axis == 0) uf = 0;
        else vf = 0;
      } else {
        g[f.axis] = 0;
        g[2 + f.axis] = 0;
        uf += g[0] * geo.ex + g[1] * geo.ey;
        vf += g[2] * geo.ex + g[3] * geo.ey;
      }
    }
    double fl[4];
    viscous_flux(mu, g, f.axis == 0 ? 1 : 0, f.axis == 1 ? 1 : 0, uf, vf, fl);
    for (int v = 0; v < 4; ++v) {
      a.face_flux[v * nf + q] = fl[v] * f.length;
      if (!std::isfinite(fl[v])) bad = 1;
    }
    fc[q] = (4.0 / 3) * nu / distance;
  }
#pragma omp parallel for reduction(| : bad)
  for (size_t q = 0; q < nw; ++q) {
    const amr_wall& wall = m.walls[q];
    const amr_wall_sample& sample = m.wall_samples[q];
    const int k = sample.donor, owner = static_cast<int>(wall.owner_cell);
    double g[4], fl[4];
    wall_gradient(wall.nx, wall.ny, sample.distance, -w[n + k], -w[2 * n + k], g);
    viscous_flux(a.mu[k], g, wall.nx, wall.ny, 0, 0, fl);
    for (int v = 0; v < 4; ++v) {
      a.wall_flux[v * nw + q] = fl[v] * wall.length;
      if (!std::isfinite(fl[v])) bad = 1;
    }
    wc[q] = (4.0 / 3) * std::max(a.mu[k] / w[k], a.mu[owner] / w[owner]) / sample.distance;
  }
  if (bad) return 1;
#pragma omp parallel for
  for (size_t k = 0; k < n; ++k) {
    const amr_cell& c = m.cells[k];
    double sum[4] = {}, spectral = 0;
    for (size_t pos = m.face_offsets[k]; pos < m.face_offsets[k + 1]; ++pos) {
      const size_t q = m.cell_faces[pos];
      const double sign = m.faces[q].left == static_cast<int>(k) ? 1 : -1;
      for (int v = 0; v < 4; ++v) sum[v] += sign * a.face_flux[v * nf + q];
      spectral += fc[q] * m.faces[q].length;
    }
    for (size_t q = c.wall_begin; q < c.wall_begin + c.wall_count; ++q) {
      for (int v = 0; v < 4; ++v) sum[v] += a.wall_flux[v * nw + q];
      spectral += wc[q] * m.walls[q].length;
    }
    for (int v = 0; v < 4; ++v) a.residual[v * n + k] = sum[v] / c.volume;
    a.cell_dt[k] = spectral > 0 ? t.viscous_cfl * c.volume / spectral : INFINITY;
  }
  if (t.stabilization == VISCOUS_SRD && m.small_cells) {
    // The diffusion rate goes through the same SRD map as the state.
    Vec rates(n), groups(n);
    for (size_t k = 0; k < n; ++k) rates[k] = t.viscous_cfl / a.cell_dt[k];
    for (size_t g = 0; g < n; ++g) {
      double sum = 0;
      for (size_t pos = m.group_offsets[g]; pos < m.group_offsets[g + 1]; ++pos) {
        const size_t k = m.group_members[pos];
        sum += m.cells[k].volume * rates[k] / csr_count(m.member_offsets, k);
      }
      groups[g] = sum / m.group_volume[g];
    }
    for (size_t k = 0; k < n; ++k) {
      double sum = 0;
      for (size_t pos = m.member_offsets[k]; pos < m.member_offsets[k + 1]; ++pos) sum += groups[m.member_groups[pos]];
      const double rate = sum / csr_count(m.member_offsets, k);
      a.cell_dt[k] = rate > 0 ? t.viscous_cfl / rate : INFINITY;
    }
  }
  return 0;
}

struct SaFields {
  Vec gradient, density_gradient, diffusion, source, source_rate, face_flux, wall_flux, residual, cell_dt;
};

int sa_evaluate(const amr_mesh_view& m, const amr_params& p, const Vec& u, const Vec& w, const ViscousFields& mean,
                const Vec& meanflux, SaFields* out) {
  const size_t n = m.n_cells, nf = m.n_faces, nw = m.n_walls;
  const amr_transport& t = m.transport;
  if (!out || u.size() != 5 * n || !m.wall_distance) return 1;
  SaFields& a = *out;
  a.gradient.assign(2 * n, 0);
  a.density_gradient.assign(2 * n, 0);
  a.diffusion.resize(n);
  a.source.resize(n);
  a.source_rate.resize(n);
  a.face_flux.resize(nf);
  a.wall_flux.resize(nw);
  a.residual.resize(n);
  a.cell_dt.resize(n);
  Vec phi(n), fc(nf), wc(nw);
  for (size_t k = 0; k < n; ++k) {
    phi[k] = u[4 * n + k] / w[k];
    if (!std::isfinite(phi[k])) return 1;
  }
  for (size_t k = 0; k < n; ++k) {
    for (size_t q = m.visc_offsets[k]; q < m.visc_offsets[k + 1]; ++q) {
      const amr_grad_term term = m.visc_terms[q];
      const double ph = term.source >= 0 ? phi[term.source] : 0, rho = term.source >= 0 ? w[term.source] : w[k];
      a.gradient[k] += term.x * (ph - phi[k]);
      a.gradient[n + k] += term.y * (ph - phi[k]);
      a.density_gradient[k] += term.x * (rho - w[k]);
      a.density_gradient[n + k] += term.y * (rho - w[k]);
    }
    const SaPoint model = sa_point(phi[k], mean.mu[k] / w[k], m.wall_distance[k],
                                   std::fabs(mean.gradient[2 * n + k] - mean.gradient[n + k]));
    a.diffusion[k] = model.diffusion;
    a.source_rate[k] = model.source_rate;
    a.source[k] = sa_source(model, w[k], a.gradient[k], a.gradient[n + k], a.density_gradient[k], a.density_gradient[n + k]);
    if (!(a.diffusion[k] > 0) || !std::isfinite(a.source[k]) || !std::isfinite(a.source_rate[k])) return 1;
  }
  for (size_t q = 0; q < nf; ++q) {
    const amr_face f = m.faces[q];
    const amr_face_geometry geo = m.face_geometry[q];
    const int inside = f.left >= 0 ? f.left : f.right;
    const double mass = meanflux[q];
    const double left = f.left >= 0 ? phi[f.left] : t.sa_nu_inf, right = f.right >= 0 ? phi[f.right] : t.sa_nu_inf;
    double diff = 0, coefficient = 0;
    if (f.boundary < 0) {
      const int l = f.left, r = f.right;
      const double gl[4] = {a.gradient[l], a.gradient[n + l], 0, 0}, gr[4] = {a.gradient[r], a.gradient[n + r], 0, 0};
      double g[4];
      face_gradient(gl, gr, geo.dx, geo.dy, phi[r] - phi[l], 0, g);
      coefficient = .5 * (w[l] * a.diffusion[l] + w[r] * a.diffusion[r]);
      diff = coefficient * g[f.axis] * f.length;
      fc[q] = coefficient / geo.distance * f.length;
    } else {
      const int bc = side_kind(m, p, f.boundary);
      const bool inflow = f.left < 0 ? mass > 0 : mass < 0;
      const double sign = f.boundary % 2 ? 1 : -1;
      if (bc == BC_NOSLIP || inflow) {
        const double wall = bc == BC_NOSLIP ? 0 : t.sa_nu_inf;
        coefficient = bc == BC_NOSLIP ? mean.mu[inside] / sigma : w[inside] * a.diffusion[inside];
        diff = sign * coefficient * (wall - phi[inside]) / m.face_samples[q].distance * f.length;
        fc[q] = coefficient / m.face_samples[q].distance * f.length;
      } else {
        fc[q] = 0;
      }
    }
    a.face_flux[q] = mass * (mass >= 0 ? left : right) - diff;
  }
  for (size_t q = 0; q < nw; ++q) {
    const amr_wall_sample sample = m.wall_samples[q];
    const amr_wall wall = m.walls[q];
    const int k = sample.donor;
    wc[q] = mean.mu[k] / sigma / sample.distance * wall.length;
    a.wall_flux[q] = -wc[q] * phi[k];
  }
  for (size_t k = 0; k < n; ++k) {
    const amr_cell c = m.cells[k];
    double net = 0, spectral = 0;
    for (size_t pos = m.face_offsets[k]; pos < m.face_offsets[k + 1]; ++pos) {
      const size_t q = m.cell_faces[pos];
      net += (m.faces[q].left == static_cast<int>(k) ? 1 : -1) * a.face_flux[q];
      spectral += fc[q];
    }
    for (size_t q = c.wall_begin; q < c.wall_begin + c.wall_count; ++q) {
      net -= a.wall_flux[q];
      spectral += wc[q];
    }
    a.residual[k] = -net / c.volume + a.source[k];
    const double rate = 2 * spectral / (w[k] * c.volume) + a.source_rate[k];
    a.cell_dt[k] = rate > 0 ? t.sa_cfl / rate : INFINITY;
    if (!std::isfinite(a.residual[k])) return 1;
  }
  return 0;
}

// Full right-hand side L(u); viscous_dt (optional) gets the viscous/SA local time step limit.
bool residual(const amr_mesh_view& m, const amr_params& p, const Vec& u, Vec* r, Vec* w, Vec* flux, Vec* speed,
              Vec* viscous_dt = nullptr) {
  const size_t n = m.n_cells, nf = m.n_faces;
  if (!primitive(u, n, p.gamma, w)) return false;
  flux->resize(4 * nf);
  speed->resize(nf);
  r->resize(4 * n);
  int bad = 0;
#pragma omp parallel for reduction(| : bad)
  for (size_t q = 0; q < nf; ++q) {
    const amr_face& f = m.faces[q];
    const int inside = f.left >= 0 ? f.left : f.right;
    double a[4], b[4];
    for (int v = 0; v < 4; ++v) {
      a[v] = (*w)[v * n + (f.left >= 0 ? f.left : inside)];
      b[v] = (*w)[v * n + (f.right >= 0 ? f.right : inside)];
    }
    if (f.boundary >= 0) {
      double* outside = f.left < 0 ? a : b;
      const double* interior = f.left < 0 ? b : a;
      if (!boundary_state(m.boundaries[f.boundary], p.bc[f.boundary], p.gamma, f.boundary, interior, outside)) {
        bad = 1;
        continue;
      }
    }
    const int normal = 1 + f.axis, tangent = 2 - f.axis;
    double local[4];
    const int bc = f.boundary < 0 ? -1 : side_kind(m, p, f.boundary);
    if (bc == BC_NOSLIP) {
      local[0] = 0;
      local[1] = (*w)[3 * n + inside];
      local[2] = 0;
      local[3] = 0;
    } else if (p.flux == FLUX_HLLC) {
      hllc(a[0], a[normal], a[tangent], a[3], b[0], b[normal], b[tangent], b[3], p.gamma, local);
    } else {
      ausm(a[0], a[normal], a[tangent], a[3], b[0], b[normal], b[tangent], b[3], p.gamma, p.ausm_mco, local);
    }
    (*flux)[q] = local[0] * f.length;
    (*flux)[normal * nf + q] = local[1] * f.length;
    (*flux)[tangent * nf + q] = local[2] * f.length;
    (*flux)[3 * nf + q] = local[3] * f.length;
    (*speed)[q] = std::max(std::fabs(a[normal]) + std::sqrt(p.gamma * a[3] / a[0]),
                           std::fabs(b[normal]) + std::sqrt(p.gamma * b[3] / b[0]));
    for (double value : local)
      if (!std::isfinite(value)) bad = 1;
  }
  if (bad) return false;
#pragma omp parallel for
  for (size_t k = 0; k < n; ++k) {
    const amr_cell& c = m.cells[k];
    double sum[4] = {};
    for (size_t q = m.face_offsets[k]; q < m.face_offsets[k + 1]; ++q) {
      const size_t face = m.cell_faces[q];
      const double sign = m.faces[face].left == static_cast<int>(k) ? 1.0 : -1.0;
      for (int v = 0; v < 4; ++v) sum[v] += sign * (*flux)[v * nf + face];
    }
    for (size_t q = c.wall_begin; q < c.wall_begin + c.wall_count; ++q) {
      const amr_wall& wall = m.walls[q];
      sum[1] += (*w)[3 * n + k] * wall.nx * wall.length;
      sum[2] += (*w)[3 * n + k] * wall.ny * wall.length;
    }
    for (int v = 0; v < 4; ++v) (*r)[v * n + k] = -sum[v] / c.volume;
  }
  if (m.transport.model == MODEL_LAMINAR) {
    ViscousFields fields;
    if (viscous_evaluate(m, p, *w, &fields, &u)) return false;
    for (size_t q = 0; q < r->size(); ++q) (*r)[q] += fields.residual[q];
    for (size_t q = 0; q < flux->size(); ++q) (*flux)[q] -= fields.face_flux[q];
    if (m.transport.sa_model) {
      SaFields sa;
      if (sa_evaluate(m, p, u, *w, fields, *flux, &sa)) return false;
      r->resize(5 * n);
      for (size_t k = 0; k < n; ++k) {
        (*r)[4 * n + k] = sa.residual[k];
        fields.cell_dt[k] = std::min(fields.cell_dt[k], sa.cell_dt[k]);
      }
    }
    if (viscous_dt) *viscous_dt = std::move(fields.cell_dt);
  }
  return true;
}

// Zero-degree state redistribution over the SRD neighbourhoods (only when small cells exist).
void redistribute(const amr_mesh_view& m, Vec* u, Vec* groups, Vec* out) {
  if (m.small_cells == 0) return;
  const size_t n = m.n_cells;
  groups->assign(4 * n, 0);
  out->resize(4 * n);
#pragma omp parallel for
  for (size_t g = 0; g < n; ++g)
    for (int v = 0; v < 4; ++v) {
      double sum = 0;
      for (size_t q = m.group_offsets[g]; q < m.group_offsets[g + 1]; ++q) {
        const size_t k = m.group_members[q];
        sum += m.cells[k].volume * (*u)[v * n + k] / csr_count(m.member_offsets, k);
      }
      (*groups)[v * n + g] = sum / m.group_volume[g];
    }
#pragma omp parallel for
  for (size_t k = 0; k < n; ++k)
    for (int v = 0; v < 4; ++v) {
      double sum = 0;
      for (size_t q = m.member_offsets[k]; q < m.member_offsets[k + 1]; ++q) sum += (*groups)[v * n + m.member_groups[q]];
      (*out)[v * n + k] = sum / csr_count(m.member_offsets, k);
    }
  u->swap(*out);
}

bool viscous_ready(const amr_mesh_view& m) {
  const amr_transport& t = m.transport;
  if (t.model == MODEL_EULER)
    for (const amr_boundary& b : m.boundaries)
      if (b.kind == BC_NOSLIP) return false;
  if (t.sa_model && !m.wall_distance) return false;
  if (t.model == MODEL_LAMINAR && (!m.visc_offsets || (m.n_walls && !m.wall_samples) || !m.face_samples || !m.face_geometry))
    return false;
  return true;
}

std::string run(const amr_mesh_view& m, const amr_params& p, double tend, long maxsteps, Vec* state,
                amr_state_meta* s, long* status) {
  const size_t n = m.n_cells;
  const size_t nv = m.transport.sa_model ? 5 : 4;
  *status = 1;
  if (!viscous_ready(m) || n == 0 || state->size() != nv * n || !std::isfinite(s->time) || s->time < 0 ||
      !std::isfinite(tend) || tend < s->time || maxsteps < 1 || !(p.gamma > 1) || !std::isfinite(p.gamma) ||
      !(p.cfl > 0) || p.cfl > 0.5 || !std::isfinite(p.dt_fixed) || p.dt_fixed < 0 || !std::isfinite(p.ausm_mco) ||
      p.ausm_mco <= 0 || p.time_integrator != 0 || p.limiter != 3 || (p.flux != FLUX_HLLC && p.flux != FLUX_AUSMPUP))
    return "static AMR requires first-order RK3, valid state/options and CFL <= 0.5";
  for (int side = 0; side < 4; ++side)
    if (!boundary_valid(m.boundaries[side], p.bc[side], side)) return "invalid AMR boundary configuration";
  Vec& u = *state;
  Vec r, w, flux, speed, u1(nv * n), u2(nv * n), groups, out, base_r, viscous_dt;
  while (s->time < tend) {
    if (s->steps >= maxsteps) return "maxsteps reached before tend";
    if (!residual(m, p, u, &r, &w, &flux, &speed, &viscous_dt)) return "nonphysical base state or nonfinite face flux";
    double dt = p.dt_fixed > 0 ? p.dt_fixed : std::numeric_limits<double>::infinity();
    if (p.dt_fixed == 0)
      for (size_t k = 0; k < n; ++k) {
        const amr_cell& c = m.cells[k];
        double spectral = 0;
        for (size_t q = m.face_offsets[k]; q < m.face_offsets[k + 1]; ++q) {
          const size_t f = m.cell_faces[q];
          spectral += speed[f] * m.faces[f].length;
        }
        const double a = std::sqrt(p.gamma * w[3 * n + k] / w[k]);
        for (size_t q = c.wall_begin; q < c.wall_begin + c.wall_count; ++q) spectral += a * m.walls[q].length;
        if (spectral > 0) dt = std::min(dt, p.cfl * (m.agglomerate_small ? c.volume : (c.x1 - c.x0) * (c.y1 - c.y0)) / spectral);
      }
    if (p.dt_fixed == 0 && m.transport.model == MODEL_LAMINAR)
      for (double value : viscous_dt) dt = std::min(dt, value);
    const double remaining = tend - s->time;
    dt = std::min(dt, remaining);
    const double time_tol = 32 * std::numeric_limits<double>::epsilon() * std::max({std::fabs(tend), std::fabs(s->time), dt});
    if (remaining - dt <= time_tol) dt = remaining;
    base_r = r;
    bool accepted = false;
    for (int attempt = 0; attempt < 21; ++attempt) {
      if (!std::isfinite(dt) || !(dt > 0) || s->time + dt == s->time) return "physical time step underflow";
      for (size_t q = 0; q < u.size(); ++q) u1[q] = u[q] + dt * base_r[q];
      redistribute(m, &u1, &groups, &out);
      bool ok = residual(m, p, u1, &r, &w, &flux, &speed);
      if (ok) {
        for (size_t q = 0; q < u.size(); ++q) u2[q] = 0.75 * u[q] + 0.25 * (u1[q] + dt * r[q]);
        redistribute(m, &u2, &groups, &out);
        ok = residual(m, p, u2, &r, &w, &flux, &speed);
      }
      if (ok) {
        for (size_t q = 0; q < u.size(); ++q) u1[q] = u[q] / 3 + (2.0 / 3) * (u2[q] + dt * r[q]);
        redistribute(m, &u1, &groups, &out);
        ok = primitive(u1, n, p.gamma, &w);
        if (m.transport.sa_model)
          for (size_t k = 0; k < n; ++k) ok = ok && std::isfinite(u1[4 * n + k]);
      }
      if (ok) {
        u.swap(u1);
        s->time = dt == remaining ? tend : s->time + dt;
        ++s->steps;
        accepted = true;
        break;
      }
      ++s->rejected;
      dt *= 0.5;
    }
    if (!accepted) return "RK candidates failed admissibility after 21 trials; physical step not committed";
  }
  *status = 0;
  return "";
}

int report(char* err, size_t len, const std::string& message) {
  if (err && len) {
    const size_t k = std::min(message.size(), len - 1);
    std::memcpy(err, message.data(), k);
    err[k] = 0;
  }
  return 1;
}

}  // namespace oracle

using namespace oracle;

extern "C" {

// static_amr2d_rhs: r gets nv * n values.  0 ok, 1 nonphysical state / invalid input.
int oracle_rhs(const amr_mesh_view* m, const amr_params* p, const double* u, uint64_t u_len, double* r, uint64_t r_len) {
  if (!m || !p || !u || !r || !viscous_ready(*m) || m->n_cells == 0) return 1;
  const size_t nv = m->transport.sa_model ? 5 : 4;
  if (u_len != nv * m->n_cells || r_len != nv * m->n_cells || !(p->gamma > 1)) return 1;
  for (int side = 0; side < 4; ++side)
    if (!boundary_valid(m->boundaries[side], p->bc[side], side)) return 1;
  Vec state(u, u + u_len), out, w, flux, speed;
  if (!residual(*m, *p, state, &out, &w, &flux, &speed)) return 1;
  std::memcpy(r, out.data(), r_len * sizeof(double));
  return 0;
}

// static_amr2d_run: advances u (nv * n, in place) to tend.  0 ok, 1 failure (message in err).
int oracle_run(const amr_mesh_view* m, const amr_params* p, double tend, int64_t maxsteps, double* u, uint64_t u_len,
               amr_state_meta* meta, char* err, size_t err_len) {
  if (!m || !p || !u || !meta) return report(err, err_len, "null argument");
  Vec state(u, u + u_len);
  long status = 1;
  const std::string message = run(*m, *p, tend, maxsteps, &state, meta, &status);
  if (state.size() == u_len) std::memcpy(u, state.data(), u_len * sizeof(double));
  if (status) return report(err, err_len, message);
  return 0;
}

int oracle_redistribute(const amr_mesh_view* m, double* u, uint64_t u_len) {
  if (!m || !u) return 1;
  const size_t nv = m->transport.sa_model ? 5 : 4;
  if (u_len != nv * m->n_cells) return 1;
  Vec state(u, u + u_len), groups, out;
  redistribute(*m, &state, &groups, &out);
  std::memcpy(u, state.data(), std::min<size_t>(u_len, state.size()) * sizeof(double));
  return 0;
}

// Steady monitor at a state (src/quick_airfoil.cpp:75-85): out[0..4] raw residual norms
// (SA divided by nutilde_inf), out[5..7] CL, CD, CM; out[8..10] fx, fy, mz.
// region = {xmin, xmax, ymin, ymax}; moment = {x, y}.  0 ok, 1 failure.
int oracle_monitor(const amr_mesh_view* m, const amr_params* p, const double* u, uint64_t u_len, const double* region,
                   const double* moment, double mach, double aoa_degrees, double* out) {
  if (!m || !p || !u || !region || !moment || !out) return 1;
  const size_t n = m->n_cells, nw = m->n_walls, nv = m->transport.sa_model ? 5 : 4;
  if (u_len != nv * n) return 1;
  std::vector<double> rhs(nv * n);
  if (oracle_rhs(m, p, u, u_len, rhs.data(), rhs.size())) return 1;
  long double sum[5] = {}, volume = 0;
  for (size_t k = 0; k < n; ++k) {
    const amr_cell& c = m->cells[k];
    if (c.cx < region[0] || c.cx > region[1] || c.cy < region[2] || c.cy > region[3]) continue;
    volume += c.volume;
    for (size_t v = 0; v < nv; ++v) sum[v] += c.volume * std::fabs(rhs[v * n + k]);
  }
  if (!(volume > 0)) return 1;
  for (int v = 0; v < 5; ++v) out[v] = v < static_cast<int>(nv) ? static_cast<double>(sum[v] / volume) : 0.0;
  if (m->transport.sa_model) out[4] /= m->transport.sa_nu_inf;
  Vec wall_flux(4 * nw, 0.0);
  if (m->transport.model == MODEL_LAMINAR) {
    Vec w(4 * n);
    for (size_t k = 0; k < n; ++k) {
      w[k] = u[k];
      w[n + k] = u[n + k] / u[k];
      w[2 * n + k] = u[2 * n + k] / u[k];
      w[3 * n + k] = (p->gamma - 1) * (u[3 * n + k] - .5 * (u[n + k] * w[n + k] + u[2 * n + k] * w[2 * n + k]));
    }
    ViscousFields fields;
    Vec state(u, u + u_len);
    if (viscous_evaluate(*m, *p, w, &fields, &state)) return 1;
    wall_flux = fields.wall_flux;
  }
  const double alpha = aoa_degrees * std::acos(-1.) / 180, cs = std::cos(alpha), sn = std::sin(alpha);
  const double qinf = .5 * mach * mach;
  double fx = 0, fy = 0, mz = 0;
  for (size_t i = 0; i < nw; ++i) {
    const amr_wall& w = m->walls[i];
    const size_t k = w.owner_cell;
    const double pressure = (p->gamma - 1) * (u[3 * n + k] - .5 * (u[n + k] * u[n + k] + u[2 * n + k] * u[2 * n + k]) / u[k]);
    const double a = pressure * w.nx * w.length - wall_flux[nw + i], b = pressure * w.ny * w.length - wall_flux[2 * nw + i];
    fx += a;
    fy += b;
    mz += ((w.x0 + w.x1) * .5 - moment[0]) * b - ((w.y0 + w.y1) * .5 - moment[1]) * a;
  }
  out[5] = (-fx * sn + fy * cs) / qinf;
  out[6] = (fx * cs + fy * sn) / qinf;
  out[7] = mz / qinf;
  out[8] = fx;
  out[9] = fy;
  out[10] = mz;
  return 0;
}

// Point functions for known-answer tests.
void oracle_flux(int kind, const double* left, const double* right, double gamma, double mco, double* F) {
  if (kind == FLUX_HLLC) hllc(left[0], left[1], left[2], left[3], right[0], right[1], right[2], right[3], gamma, F);
  else if (kind == FLUX_AUSMPUP) ausm(left[0], left[1], left[2], left[3], right[0], right[1], right[2], right[3], gamma, mco, F);
  else exact_flux(left[0], left[1], left[2], left[3], gamma, F);
}
int oracle_boundary_state(const amr_boundary* b, int legacy, double gamma, int side, const double* wi, double* wb) {
  return boundary_state(*b, legacy, gamma, side, wi, wb) ? 1 : 0;
}
int oracle_boundary_valid(const amr_boundary* b, int legacy, int side) { return boundary_valid(*b, legacy, side) ? 1 : 0; }
// out = {fv1, modified_vorticity, fw, source, source_rate, diffusion, eddy_kinematic}
void oracle_sa_point(double working, double molecular, double distance, double omega, double* out) {
  const SaPoint a = sa_point(working, molecular, distance, omega);
  out[0] = a.fv1;
  out[1] = a.modified_vorticity;
  out[2] = a.fw;
  out[3] = a.source;
  out[4] = a.source_rate;
  out[5] = a.diffusion;
  out[6] = sa_eddy(working, molecular);
}
double oracle_viscosity(const amr_transport* t, double temperature) { return viscosity(*t, temperature); }

// Seconds of one residual evaluation averaged over `repeats` (CPU baseline timing).
double oracle_time_rhs(const amr_mesh_view* m, const amr_params* p, const double* u, uint64_t u_len, int repeats) {
  Vec state(u, u + u_len), out, w, flux, speed;
  const auto start = std::chrono::steady_clock::now();
  for (int i = 0; i < repeats; ++i) residual(*m, *p, state, &out, &w, &flux, &speed);
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count() / repeats;
}

}  // extern "C"
