// amr_physics.h -- point physics of the cut-cell time step for gfx950, templated on the
// arithmetic type (double for the FP64 path, float for the FP32 path).
//
// Same mathematics as the reference's shared headers, written for the device:
//   HLLC (Toro-Spruce-Speares, Einfeldt/Roe speeds) .......... src/flux.h:42-103
//   AUSM+-up (Liou 2006; Kp .25, Ku .75, sigma 1, beta 1/8) .. src/flux.h:121-176
//   characteristic boundary states ........................... src/static_amr2d_boundary.h:50-102
//   Sutherland viscosity, Newtonian stress, face and wall
//   gradient reconstructions ................................. src/static_amr2d_viscous.h:157-179
//   NASA-TMR SA-neg point model (ft2, Note 1c) ............... src/sa_model.h:12-39
// Operation order is this file's own (FMA contraction is on); results agree with the FP64
// oracle to rounding, which the parity tests bound.
#ifndef HIPAMR_AMR_PHYSICS_H
#define HIPAMR_AMR_PHYSICS_H

#include <hip/hip_runtime.h>

namespace hipamr {

enum : int {
  KIND_TRANSMISSIVE = 1, KIND_REFLECT = 2, KIND_FARFIELD = 3, KIND_TOTAL_INLET = 4,
  KIND_PRESSURE_OUTLET = 5, KIND_NOSLIP = 6
};

template <class T> struct Tiny;
template <> struct Tiny<double> { static constexpr double value = 1e-300; };
template <> struct Tiny<float> { static constexpr float value = 1e-30f; };

// One side of the domain, resolved kind (the -1 "inherit" case is folded in at upload).
template <class T>
struct Side {
  int kind;
  T reference[4];  // rho, u, v, p
  T total_pressure, total_temperature, gas_constant;
  T direction[2];
  T outlet_pressure;
  T wall_velocity[2];
};

// Primitive state (rho, u, v, p) held in registers; velocities in global x/y components.
template <class T>
struct Prim {
  T r, u, v, p;
};

template <class T>
__device__ __forceinline__ bool admissible(const Prim<T>& w) {
  return isfinite(w.r) && w.r > T(0) && isfinite(w.u) && isfinite(w.v) && isfinite(w.p) && w.p > T(0);
}

// Euler flux of one state in the face frame (normal velocity un, tangential ut).
template <class T>
__device__ __forceinline__ void euler_flux(T r, T un, T ut, T p, T E, T* f) {
  const T mass = r * un;
  f[0] = mass;
  f[1] = mass * un + p;
  f[2] = mass * ut;
  f[3] = un * (E + p);
}

template <class T>
__device__ __forceinline__ void hllc_flux(T rL, T uL, T vL, T pL, T rR, T uR, T vR, T pR, T g, T* f) {
  const T gm1 = g - T(1);
  const T iL = T(1) / rL, iR = T(1) / rR;
  const T aL = sqrt(g * pL * iL), aR = sqrt(g * pR * iR);
  const T EL = pL / gm1 + T(0.5) * rL * (uL * uL + vL * vL);
  const T ER = pR / gm1 + T(0.5) * rR * (uR * uR + vR * vR);
  const T sL = sqrt(rL), sR = sqrt(rR), iS = T(1) / (sL + sR);
  const T HL = (EL + pL) * iL, HR = (ER + pR) * iR;
  const T uT = (sL * uL + sR * uR) * iS, vT = (sL * vL + sR * vR) * iS;
  const T HT = (sL * HL + sR * HR) * iS;
  const T aT = sqrt(fmax(gm1 * (HT - T(0.5) * (uT * uT + vT * vT)), Tiny<T>::value));
  const T SL = fmin(uL - aL, uT - aT), SR = fmax(uR + aR, uT + aT);
  if (SL >= T(0)) {
    euler_flux(rL, uL, vL, pL, EL, f);
    return;
  }
  if (SR <= T(0)) {
    euler_flux(rR, uR, vR, pR, ER, f);
    return;
  }
  const T mL = rL * (SL - uL), mR = rR * (SR - uR);
  const T Ss = (pR - pL + uL * mL - uR * mR) / (mL - mR);
  // Star state of the upwind side K: U*K = c (1, S*, vK, EK/rK + (S*-uK)(S* + pK/mK)).
  const bool left = Ss >= T(0);
  const T r = left ? rL : rR, u = left ? uL : uR, v = left ? vL : vR, p = left ? pL : pR;
  const T E = left ? EL : ER, S = left ? SL : SR, m = left ? mL : mR;
  const T c = m / (S - Ss);
  T fk[4];
  euler_flux(r, u, v, p, E, fk);
  f[0] = fk[0] + S * (c - r);
  f[1] = fk[1] + S * (c * Ss - r * u);
  f[2] = fk[2] + S * (c * v - r * v);
  f[3] = fk[3] + S * (c * (E / r + (Ss - u) * (Ss + p / m)) - E);
}

template <class T>
__device__ __forceinline__ T split_m2(T M, T s) { return s * T(0.25) * (M + s) * (M + s); }
template <class T>
__device__ __forceinline__ T split_m4(T M, T s) {
  if (fabs(M) >= T(1)) return T(0.5) * (M + s * fabs(M));
  return split_m2(M, s) * (T(1) - s * T(2) * split_m2(M, -s));  // 16 beta = 2
}
template <class T>
__device__ __forceinline__ T split_p5(T M, T s, T alpha) {
  if (fabs(M) >= T(1)) return T(0.5) * (T(1) + s * (M > T(0) ? T(1) : (M < T(0) ? T(-1) : T(0))));
  return split_m2(M, s) * ((s * T(2) - M) - s * T(16) * alpha * M * split_m2(M, -s));
}

template <class T>
__device__ __forceinline__ void ausm_flux(T rL, T uL, T vL, T pL, T rR, T uR, T vR, T pR, T g, T mco, T* f) {
  const T a12 = T(0.5) * (sqrt(g * pL / rL) + sqrt(g * pR / rR));
  const T ia = T(1) / a12;
  const T rho12 = T(0.5) * (rL + rR), ML = uL * ia, MR = uR * ia;
  const T mb2 = T(0.5) * (uL * uL + uR * uR) * ia * ia;
  const T m02 = fmin(T(1), fmax(mb2, mco * mco));
  const T m0 = sqrt(m02);
  const T fa = m0 * (T(2) - m0);
  const T alpha = T(3.0 / 16.0) * (T(-4) + T(5) * fa * fa);
  const T m12 = split_m4(ML, T(1)) + split_m4(MR, T(-1)) -
                (T(0.25) / fa) * fmax(T(1) - mb2, T(0)) * (pR - pL) / (rho12 * a12 * a12);
  const T pp = split_p5(ML, T(1), alpha), pm = split_p5(MR, T(-1), alpha);
  const T p12 = pp * pL + pm * pR - T(0.75) * pp * pm * (rL + rR) * fa * a12 * (uR - uL);
  const bool up = m12 > T(0);
  const T r = up ? rL : rR, u = up ? uL : uR, v = up ? vL : vR, p = up ? pL : pR;
  const T mdot = a12 * m12 * r;
  const T E = p / (g - T(1)) + T(0.5) * r * (u * u + v * v);
  f[0] = mdot;
  f[1] = mdot * u + p12;
  f[2] = mdot * v;
  f[3] = mdot * (E + p) / r;
}

// Exterior Riemann state of domain side `side` for the interior state wi.  Works on register
// values only: normal/tangent components are picked with selects, never by indexing, so the
// compiler keeps everything in VGPRs (no private-memory arrays, no generic pointers).
// Returns false on a physical/configuration failure (no clipping, no fallback).
template <class T>
__device__ __forceinline__ bool side_state(const Side<T>& b, T g, int side, const Prim<T>& wi, Prim<T>& wb) {
  const bool xaxis = (side >> 1) == 0;
  const T sign = (side & 1) ? T(1) : T(-1), gm = g - T(1);
  wb = wi;
  const int kind = b.kind;
  if (kind == KIND_TRANSMISSIVE) return true;
  if (kind == KIND_NOSLIP) {
    wb.u = T(2) * b.wall_velocity[0] - wi.u;
    wb.v = T(2) * b.wall_velocity[1] - wi.v;
    return true;
  }
  if (kind == KIND_REFLECT) {
    if (xaxis) wb.u = -wi.u;
    else wb.v = -wi.v;
    return true;
  }
  if (!admissible(wi) || !(g > T(1))) return false;
  const T wi_n = xaxis ? wi.u : wi.v, wi_t = xaxis ? wi.v : wi.u;
  const T ai = sqrt(g * wi.p / wi.r), un = sign * wi_n;
  if (kind == KIND_TOTAL_INLET) {
    if (!(un < T(0) && un > -ai)) return false;
    const T dn = sign * (xaxis ? b.direction[0] : b.direction[1]), A = g * b.gas_constant * b.total_temperature;
    if (!(dn < T(0)) || !isfinite(A)) return false;
    const T rp = un + T(2) * ai / gm, upper = sqrt(T(2) * A / (gm + T(2) * dn * dn));
    const T f0 = T(2) * sqrt(A) / gm - rp;
    const T f1 = dn * upper + T(2) * sqrt(A - T(0.5) * gm * upper * upper) / gm - rp;
    if (!(f0 >= T(0) && f1 <= T(0))) return false;
    T lo = T(0), hi = upper;
    for (int it = 0; it < 64; ++it) {  // bisection on the outgoing Riemann invariant
      const T v = T(0.5) * (lo + hi), a = sqrt(A - T(0.5) * gm * v * v);
      if (dn * v + T(2) * a / gm > rp) lo = v;
      else hi = v;
    }
    const T v = T(0.5) * (lo + hi), temp = (A - T(0.5) * gm * v * v) / (g * b.gas_constant);
    const T pressure = b.total_pressure * pow(temp / b.total_temperature, g / gm);
    wb.r = pressure / (b.gas_constant * temp);
    wb.u = v * b.direction[0];
    wb.v = v * b.direction[1];
    wb.p = pressure;
    return admissible(wb);
  }
  T out_n = wi_n, out_t = wi_t;
  if (kind == KIND_PRESSURE_OUTLET && un >= T(0)) {
    if (un >= ai) return true;  // supersonic outflow imposes nothing
    const T density = wi.r * pow(b.outlet_pressure / wi.p, T(1) / g);
    const T ab = sqrt(g * b.outlet_pressure / density), ub = un + T(2) * (ai - ab) / gm;
    if (ub >= T(0)) {
      out_n = sign * ub;
      wb.r = density;
      wb.p = b.outlet_pressure;
      wb.u = xaxis ? out_n : out_t;
      wb.v = xaxis ? out_t : out_n;
      return admissible(wb);
    }
  }
  if (kind != KIND_FARFIELD && kind != KIND_PRESSURE_OUTLET) return false;
  if (un >= ai) return true;
  const T rr = b.reference[0], ru = b.reference[1], rv = b.reference[2], rpr = b.reference[3];
  if (un <= -ai) {
    wb.r = rr;
    wb.u = ru;
    wb.v = rv;
    wb.p = rpr;
    return admissible(wb);
  }
  const T ar = sqrt(g * rpr / rr);
  const T r_n = xaxis ? ru : rv, r_t = xaxis ? rv : ru;
  const T rp = un + T(2) * ai / gm, rm = sign * r_n - T(2) * ar / gm;
  const T ab = T(0.25) * gm * (rp - rm), ub = T(0.5) * (rp + rm);
  if (!(ab > T(0)) || !isfinite(ab)) return false;
  const bool out = un >= T(0);
  const T advrho = out ? wi.r : rr, advp = out ? wi.p : rpr;
  out_t = out ? wi_t : r_t;
  out_n = sign * ub;
  const T entropy = advp / pow(advrho, g);
  wb.r = pow(ab * ab / (g * entropy), T(1) / gm);
  wb.p = entropy * pow(wb.r, g);
  wb.u = xaxis ? out_n : out_t;
  wb.v = xaxis ? out_t : out_n;
  return admissible(wb);
}

// Transport: constant or Sutherland viscosity; -1 flags a nonphysical temperature.
template <class T>
struct Transport {
  int law;  // 0 constant, 1 Sutherland
  T mu_ref, gas_constant, temperature_ref, sutherland, viscous_cfl, sa_nu_inf, sa_cfl;
};
template <class T>
__device__ __forceinline__ T viscosity(const Transport<T>& t, T temperature) {
  if (!(temperature > T(0)) || !isfinite(temperature)) return T(-1);
  if (t.law == 0) return t.mu_ref;
  const T ratio = temperature / t.temperature_ref;
  return t.mu_ref * ratio * sqrt(ratio) * (t.temperature_ref + t.sutherland) / (temperature + t.sutherland);
}

// Newtonian viscous flux through a unit normal (nx, ny); g = (ux, uy, vx, vy).
template <class T>
__device__ __forceinline__ void stress_flux(T mu, const T* g, T nx, T ny, T u, T v, T* f) {
  const T dil = T(2.0 / 3.0) * (g[0] + g[3]);
  const T txx = mu * (T(2) * g[0] - dil), txy = mu * (g[1] + g[2]), tyy = mu * (T(2) * g[3] - dil);
  f[1] = txx * nx + txy * ny;
  f[2] = txy * nx + tyy * ny;
  f[3] = u * f[1] + v * f[2];
}

// Face gradient: mean of the two cell gradients with its component along d replaced by the
// difference quotient (du, dv over d = (dx, dy)).
template <class T>
__device__ __forceinline__ void corrected_gradient(const T* L, const T* R, T dx, T dy, T du, T dv, T* g) {
  const T id2 = T(1) / (dx * dx + dy * dy);
  for (int q = 0; q < 4; ++q) g[q] = T(0.5) * (L[q] + R[q]);
  const T cu = (du - g[0] * dx - g[1] * dy) * id2, cv = (dv - g[2] * dx - g[3] * dy) * id2;
  g[0] += cu * dx;
  g[1] += cu * dy;
  g[2] += cv * dx;
  g[3] += cv * dy;
}

// ---- SA-neg point model ----
template <class T>
struct SaConst {
  static constexpr double cb1 = .1355, cb2 = .622, sigma = 2.0 / 3, kappa = .41;
  static constexpr double cw1 = cb1 / (kappa * kappa) + (1 + cb2) / sigma;
  static constexpr double cv1_cube = 7.1 * 7.1 * 7.1;
};

template <class T>
__device__ __forceinline__ T sa_eddy(T working, T molecular) {
  if (working <= T(0)) return T(0);
  const T chi = working / molecular, cube = chi * chi * chi;
  return working * cube / (cube + T(SaConst<T>::cv1_cube));
}

template <class T>
struct SaPoint {
  T source, source_rate, diffusion;
};

template <class T>
__device__ __forceinline__ SaPoint<T> sa_point(T working, T molecular, T distance, T omega) {
  using K = SaConst<T>;
  const T cb1 = T(K::cb1), cw1 = T(K::cw1), isig = T(1.0 / K::sigma), k2 = T(K::kappa * K::kappa);
  SaPoint<T> a;
  const T chi = working / molecular, cube = chi * chi * chi, id2 = T(1) / (distance * distance);
  if (working < T(0)) {
    const T fn = (T(16) + cube) / (T(16) - cube);
    a.diffusion = (molecular + working * fn) * isig;
    a.source = T(K::cb1 * (1 - 1.2)) * omega * working + cw1 * working * working * id2;
    a.source_rate = T(K::cb1 * .2) * omega + T(2) * cw1 * fabs(working) * id2;
    return a;
  }
  const T fv1 = cube / (cube + T(K::cv1_cube));
  const T fv2 = T(1) - chi / (T(1) + chi * fv1);
  const T sbar = working * fv2 * id2 / k2;
  const T sm = sbar >= T(-0.7) * omega ? omega + sbar : omega * ((T(-0.01) * omega - T(0.1) * sbar) / (T(-0.5) * omega - sbar));
  const T r = sm == T(0) ? T(10) : fmin(working / (sm * k2) * id2, T(10));
  const T r2 = r * r, r6 = r2 * r2 * r2;
  const T g = r + T(0.3) * (r6 - r), g2 = g * g, g6 = g2 * g2 * g2;
  const T fw = g * pow(T(65) / (g6 + T(64)), T(1.0 / 6.0));
  const T ft2 = T(1.2) * exp(T(-0.5) * chi * chi);
  const T production = cb1 * (T(1) - ft2) * sm, destruction = cw1 * fw - cb1 * ft2 / k2;
  a.source = production * working - destruction * working * working * id2;
  a.source_rate = fabs(production) + T(2) * fabs(destruction) * working * id2;
  a.diffusion = (molecular + working) * isig;
  return a;
}

}  // namespace hipamr

#endif
