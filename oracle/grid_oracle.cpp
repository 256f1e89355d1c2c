// grid_oracle.cpp -- TEST INFRASTRUCTURE ONLY.  FP64 CPU restatement of the reference's legacy
// GPU backend for the tensor-product grid (first order, regular grid, SSP-RK3), used by tests/
// as the checker of the HIP ugrid path.  Nothing in the product links or calls it.
//
// It follows the reference's CUDA kernels and host loop operation for operation, in plain loops
// over the same arrays (state with two ghost layers, primitive planes, x / y face flux planes):
//   boundary fill .............. src/euler2d_cuda.cu:153-180 (apply_bc_kernel)
//   primitives ................. src/euler2d_cuda.cu:182-196
//   legacy HLLC / AUSM+-up ..... src/euler2d_cuda.cu:71-146 (their own rounding: energy as
//                                p / (g - 1) + 0.5 r (u^2 + v^2), AUSM pressure term grouped
//                                left to right, energy flux mdot (E + p) / r)
//   x / y face fluxes .......... src/euler2d_cuda.cu:198-236
//   collect (residual) ......... src/euler2d_cuda.cu:238-258 (division by dx, dy)
//   wave-speed bound ........... src/euler2d_cuda.cu:271-287 (maximum taken on the bit
//                                patterns, as atomicMax on the bits does: a NaN bound wins)
//   RK combine, admissibility .. src/euler2d_cuda.cu:289-314
//   cut cells .................. src/euler2d_cuda.cu:182-258, 271-314 with a Cut2D: solid cells
//                                (area 0) get primitives (1, 0, 0, 1), residual 0 and are
//                                skipped by the bound and the admissibility test; closed faces
//                                carry no flux; a cut cell gathers its apertures' fluxes and the
//                                wall pressure term with its own pressure
//   SRD ........................ src/euler2d_cuda.cu:316-368, 461-468 (group averages over the
//                                members, regular cells keep their value, every member adds each
//                                neighbourhood's average -- here in ascending neighbourhood
//                                order, the CPU solver's order; the reference's device adds them
//                                with atomics in no fixed order -- then divides by its count),
//                                after each stage whose candidate is admissible, then the
//                                admissibility test again
//   run loop ................... src/euler2d_cuda.cu:605-693 (dt choice and clipping, three
//                                stages with a check after each, retry halving with the
//                                "hdt > 0" / "hdt < 0 or not finite" tests, accept copy of the
//                                interior, maxsteps counted per call)
//   backward Euler ............. src/euler2d_cuda.cu:370-447 with src/implicit2d.h:16-63 and
//                                src/implicit2d.cpp:11-69 (controls check, initial state check,
//                                physical dt from the spectral bound with the round-off rule,
//                                pseudo-time defect relaxation: residual, scaled defect norm per
//                                case, tolerance max(atol, rtol * first norm), local relaxation
//                                factor, candidate check with pseudo-CFL halving (12 retries),
//                                interior commit of every case, time advance that lands on tend)
// The reference's CPU solver (src/euler2d.cpp:249-698) differs from this backend only in
// rounding (pressure and energy groupings, multiplication by 1 / dx) and in the retry test
// for a step that underflows to zero; the tests pin this restatement against it within
// round-off, bit for bit on the boundary fill, and on closed forms.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "ugrid_hip.h"

namespace grid_oracle {

constexpr int NG = 2;
enum { BC_PERIODIC = 0, BC_TRANSMISSIVE = 1, BC_REFLECT = 2 };

struct Shape {
  int nx, ny, mx, my, ncase;
  size_t plane;
  explicit Shape(const ugrid_setup& s)
      : nx(s.nx), ny(s.ny), mx(s.nx + 2 * NG), my(s.ny + 2 * NG), ncase(s.ncase),
        plane(static_cast<size_t>(s.nx + 2 * NG) * (s.ny + 2 * NG)) {}
  size_t at(int c, int v, int i, int j) const {
    return (static_cast<size_t>(c) * 4 + v) * plane + static_cast<size_t>(j + NG) * mx + (i + NG);
  }
};

// ----------------------------------------------------------------- legacy fluxes --
void exact(double r, double u, double v, double p, double gm, double* f) {
  const double E = p / (gm - 1.0) + 0.5 * r * (u * u + v * v);
  f[0] = r * u;
  f[1] = r * u * u + p;
  f[2] = r * u * v;
  f[3] = u * (E + p);
}

void hllc(double rL, double uL, double vL, double pL, double rR, double uR, double vR, double pR, double gm,
          double* f) {
  const double aL = std::sqrt(gm * pL / rL);
  const double aR = std::sqrt(gm * pR / rR);
  const double EL = pL / (gm - 1.0) + 0.5 * rL * (uL * uL + vL * vL);
  const double ER = pR / (gm - 1.0) + 0.5 * rR * (uR * uR + vR * vR);
  const double sqL = std::sqrt(rL), sqR = std::sqrt(rR);
  const double HL = (EL + pL) / rL;
  const double HR = (ER + pR) / rR;
  const double den = sqL + sqR;
  const double uT = (sqL * uL + sqR * uR) / den;
  const double vT = (sqL * vL + sqR * vR) / den;
  const double HT = (sqL * HL + sqR * HR) / den;
  const double aT = std::sqrt(std::fmax((gm - 1.0) * (HT - 0.5 * (uT * uT + vT * vT)), 1e-300));
  const double SL = std::fmin(uL - aL, uT - aT);
  const double SR = std::fmax(uR + aR, uT + aT);
  double fL[4], fR[4];
  exact(rL, uL, vL, pL, gm, fL);
  exact(rR, uR, vR, pR, gm, fR);
  if (SL >= 0.0) {
    std::memcpy(f, fL, sizeof(fL));
    return;
  }
  if (SR <= 0.0) {
    std::memcpy(f, fR, sizeof(fR));
    return;
  }
  const double sstar = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR));
  if (sstar >= 0.0) {
    const double k = rL * (SL - uL) / (SL - sstar);
    const double star[4] = {k, k * sstar, k * vL, k * (EL / rL + (sstar - uL) * (sstar + pL / (rL * (SL - uL))))};
    const double own[4] = {rL, rL * uL, rL * vL, EL};
    for (int q = 0; q < 4; ++q) f[q] = fL[q] + SL * (star[q] - own[q]);
  } else {
    const double k = rR * (SR - uR) / (SR - sstar);
    const double star[4] = {k, k * sstar, k * vR, k * (ER / rR + (sstar - uR) * (sstar + pR / (rR * (SR - uR))))};
    const double own[4] = {rR, rR * uR, rR * vR, ER};
    for (int q = 0; q < 4; ++q) f[q] = fR[q] + SR * (star[q] - own[q]);
  }
}

double split_m1(double m, int s) { return 0.5 * (m + s * std::fabs(m)); }
double split_m2(double m, int s) { return s * 0.25 * (m + s) * (m + s); }
double split_m4(double m, int s) {
  return std::fabs(m) >= 1.0 ? split_m1(m, s) : split_m2(m, s) * (1.0 - s * 2.0 * split_m2(m, -s));
}
double split_p5(double m, int s, double alpha) {
  if (std::fabs(m) >= 1.0) {
    const double sign = m > 0.0 ? 1.0 : (m < 0.0 ? -1.0 : 0.0);
    return 0.5 * (1.0 + s * sign);
  }
  return split_m2(m, s) * ((s * 2.0 - m) - s * 16.0 * alpha * m * split_m2(m, -s));
}

void ausm(double rL, double uL, double vL, double pL, double rR, double uR, double vR, double pR, double gm,
          double mco, double* f) {
  const double a12 = 0.5 * (std::sqrt(gm * pL / rL) + std::sqrt(gm * pR / rR));
  const double rho12 = 0.5 * (rL + rR);
  const double ML = uL / a12, MR = uR / a12;
  const double mb2 = 0.5 * (uL * uL + uR * uR) / (a12 * a12);
  const double m02 = std::fmin(1.0, std::fmax(mb2, mco * mco));
  const double fa = std::sqrt(m02) * (2.0 - std::sqrt(m02));
  const double alpha = 3.0 / 16.0 * (-4.0 + 5.0 * fa * fa);
  const double m12 = split_m4(ML, 1) + split_m4(MR, -1) -
                     (0.25 / fa) * std::fmax(1.0 - mb2, 0.0) * (pR - pL) / (rho12 * a12 * a12);
  const double pp = split_p5(ML, 1, alpha), pm = split_p5(MR, -1, alpha);
  const double p12 = pp * pL + pm * pR - 0.75 * pp * pm * (rL + rR) * fa * a12 * (uR - uL);
  const bool from_left = m12 > 0.0;
  const double r = from_left ? rL : rR;
  const double u = from_left ? uL : uR;
  const double v = from_left ? vL : vR;
  const double p = from_left ? pL : pR;
  const double mdot = a12 * m12 * r;
  const double E = p / (gm - 1.0) + 0.5 * r * (u * u + v * v);
  f[0] = mdot;
  f[1] = mdot * u + p12;
  f[2] = mdot * v;
  f[3] = mdot * (E + p) / r;
}

void face(const ugrid_params& par, double mco, const double* L, const double* R, double* f) {
  if (par.flux == 1)
    ausm(L[0], L[1], L[2], L[3], R[0], R[1], R[2], R[3], par.gamma, mco, f);
  else
    hllc(L[0], L[1], L[2], L[3], R[0], R[1], R[2], R[3], par.gamma, f);
}

// ------------------------------------------------------------------- boundary fill --
void apply_bc(const ugrid_setup& s, double* U) {
  const Shape g(s);
  const int* bc = s.params.bc;
  for (int c = 0; c < g.ncase; ++c)
    for (int v = 0; v < 4; ++v)
      for (int j = -NG; j < g.ny + NG; ++j)
        for (int i = -NG; i < g.nx + NG; ++i) {
          if (i >= 0 && i < g.nx && j >= 0 && j < g.ny) continue;
          int si = i, sj = j;
          double sign = 1.0;
          if (i < 0) {
            si = bc[0] == BC_PERIODIC ? i + g.nx : bc[0] == BC_REFLECT ? -i - 1 : 0;
            if (bc[0] == BC_REFLECT && v == 1) sign = -sign;
          } else if (i >= g.nx) {
            si = bc[1] == BC_PERIODIC ? i - g.nx : bc[1] == BC_REFLECT ? 2 * g.nx - 1 - i : g.nx - 1;
            if (bc[1] == BC_REFLECT && v == 1) sign = -sign;
          }
          if (j < 0) {
            sj = bc[2] == BC_PERIODIC ? j + g.ny : bc[2] == BC_REFLECT ? -j - 1 : 0;
            if (bc[2] == BC_REFLECT && v == 2) sign = -sign;
          } else if (j >= g.ny) {
            sj = bc[3] == BC_PERIODIC ? j - g.ny : bc[3] == BC_REFLECT ? 2 * g.ny - 1 - j : g.ny - 1;
            if (bc[3] == BC_REFLECT && v == 2) sign = -sign;
          }
          U[g.at(c, v, i, j)] = sign * U[g.at(c, v, si, sj)];
        }
}

// ------------------------------------------------------------------------ residual --
// R = -(FX_e - FX_w) / dx - (FY_n - FY_s) / dy on interior cells, ghosts of U already filled.
void residual(const ugrid_setup& s, const std::vector<double>& mco, const double* U, double* R) {
  const Shape g(s);
  const double gm = s.params.gamma;
  std::vector<double> W(static_cast<size_t>(g.ncase) * 4 * g.plane);
  const ugrid_cut_view* cut = s.cut;
  for (int c = 0; c < g.ncase; ++c)
    for (size_t k = 0; k < g.plane; ++k) {
      const size_t b = static_cast<size_t>(c) * 4 * g.plane + k;
      if (cut && cut->vol[k] <= 0.0) {
        W[b] = 1.0;
        W[b + g.plane] = 0.0;
        W[b + 2 * g.plane] = 0.0;
        W[b + 3 * g.plane] = 1.0;
        continue;
      }
      const double r = U[b], mu = U[b + g.plane], mv = U[b + 2 * g.plane], E = U[b + 3 * g.plane];
      W[b] = r;
      W[b + g.plane] = mu / r;
      W[b + 2 * g.plane] = mv / r;
      W[b + 3 * g.plane] = (gm - 1.0) * (E - 0.5 * (mu * mu + mv * mv) / r);
    }
  const size_t nfx = static_cast<size_t>(g.nx + 1) * g.ny, nfy = static_cast<size_t>(g.nx) * (g.ny + 1);
  std::vector<double> FX(4 * nfx), FY(4 * nfy);
  for (int c = 0; c < g.ncase; ++c) {
    auto prim = [&](int i, int j, bool xface, double* out) {
      // (rho, normal velocity, tangential velocity, p)
      out[0] = W[g.at(c, 0, i, j)];
      out[1] = W[g.at(c, xface ? 1 : 2, i, j)];
      out[2] = W[g.at(c, xface ? 2 : 1, i, j)];
      out[3] = W[g.at(c, 3, i, j)];
    };
    double L[4], Rt[4], f[4];
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i <= g.nx; ++i) {
        const size_t q = static_cast<size_t>(j) * (g.nx + 1) + i;
        f[0] = f[1] = f[2] = f[3] = 0.0;
        if (!cut || cut->axl[q] > 0.0) {
          prim(i - 1, j, true, L);
          prim(i, j, true, Rt);
          face(s.params, mco[c], L, Rt, f);
        }
        for (int v = 0; v < 4; ++v) FX[v * nfx + q] = f[v];
      }
    for (int j = 0; j <= g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const size_t q = static_cast<size_t>(j) * g.nx + i;
        f[0] = f[1] = f[2] = f[3] = 0.0;
        if (!cut || cut->ayl[q] > 0.0) {
          prim(i, j - 1, false, L);
          prim(i, j, false, Rt);
          face(s.params, mco[c], L, Rt, f);
        }
        FY[q] = f[0];
        FY[nfy + q] = f[2];
        FY[2 * nfy + q] = f[1];
        FY[3 * nfy + q] = f[3];
      }
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const size_t e = static_cast<size_t>(j) * (g.nx + 1) + i, n = static_cast<size_t>(j) * g.nx + i;
        const size_t k = static_cast<size_t>(j + NG) * g.mx + (i + NG);
        if (cut && cut->vol[k] <= 0.0) {
          for (int v = 0; v < 4; ++v) R[g.at(c, v, i, j)] = 0.0;
          continue;
        }
        if (!cut || cut->flag[k] == 0.0) {
          for (int v = 0; v < 4; ++v)
            R[g.at(c, v, i, j)] = -(FX[v * nfx + e + 1] - FX[v * nfx + e]) / s.dx[i + NG] -
                                  (FY[v * nfy + n + g.nx] - FY[v * nfy + n]) / s.dy[j + NG];
          continue;
        }
        const double iv = 1.0 / cut->vol[k], pw = W[g.at(c, 3, i, j)];
        for (int v = 0; v < 4; ++v) {
          double r = -iv * (cut->axl[e + 1] * FX[v * nfx + e + 1] - cut->axl[e] * FX[v * nfx + e] +
                            cut->ayl[n + g.nx] * FY[v * nfy + n + g.nx] - cut->ayl[n] * FY[v * nfy + n]);
          if (v == 1) r -= iv * pw * cut->wax[k];
          if (v == 2) r -= iv * pw * cut->way[k];
          R[g.at(c, v, i, j)] = r;
        }
      }
  }
}

// ------------------------------------------------------------------------ run loop --
double bound_max(const ugrid_setup& s, const double* U, int c) {
  const Shape g(s);
  const double gm = s.params.gamma;
  uint64_t best = 0;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      if (s.cut && s.cut->vol[static_cast<size_t>(j + NG) * g.mx + (i + NG)] <= 0.0) continue;
      const size_t k = g.at(c, 0, i, j);
      const double r = U[k], mu = U[k + g.plane], mv = U[k + 2 * g.plane], E = U[k + 3 * g.plane];
      const double u = mu / r, v = mv / r;
      const double p = (gm - 1.0) * (E - 0.5 * (mu * mu + mv * mv) / r);
      const double a = std::sqrt(gm * p / r);
      const double val = (std::fabs(u) + a) / s.dx[i + NG] + (std::fabs(v) + a) / s.dy[j + NG];
      uint64_t bits;
      std::memcpy(&bits, &val, 8);
      if (bits > best) best = bits;
    }
  double out;
  std::memcpy(&out, &best, 8);
  return out;
}

bool any_bad(const ugrid_setup& s, const double* U) {
  const Shape g(s);
  const double gm = s.params.gamma;
  for (int c = 0; c < g.ncase; ++c)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        if (s.cut && s.cut->vol[static_cast<size_t>(j + NG) * g.mx + (i + NG)] <= 0.0) continue;
        const size_t k = g.at(c, 0, i, j);
        const double r = U[k], mu = U[k + g.plane], mv = U[k + 2 * g.plane], E = U[k + 3 * g.plane];
        const double p = (gm - 1.0) * (E - 0.5 * (mu * mu + mv * mv) / r);
        if (!(r > 0.0) || !(p > 0.0) || !std::isfinite(r) || !std::isfinite(p)) return true;
      }
  return false;
}

// State redistribution of every case in place (first order: the neighbourhood average).
void srd(const ugrid_setup& s, double* U) {
  const ugrid_cut_view* cut = s.cut;
  if (!cut || cut->nsmall == 0) return;
  const Shape g(s);
  const int ns = cut->nsmall;
  std::vector<double> Qh(static_cast<size_t>(ns) * 4), acc(4 * g.plane);
  for (int c = 0; c < g.ncase; ++c) {
    double* Uc = U + static_cast<size_t>(c) * 4 * g.plane;
    for (int q = 0; q < ns; ++q)
      for (int v = 0; v < 4; ++v) {
        double sum = 0.0;
        for (int m = 0; m < static_cast<int>(cut->scnt[q]); ++m) {
          const size_t r = static_cast<size_t>(cut->smem[25 * static_cast<size_t>(q) + m]);
          sum += cut->vol[r] / cut->nbrcount[r] * Uc[v * g.plane + r];
        }
        Qh[static_cast<size_t>(q) * 4 + v] = sum / cut->svhat[q];
      }
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const size_t k = static_cast<size_t>(j + NG) * g.mx + (i + NG);
        for (int v = 0; v < 4; ++v)
          acc[v * g.plane + k] = (cut->vol[k] > 0.0 && cut->smallpos[k] < 0.0) ? Uc[v * g.plane + k] : 0.0;
      }
    for (int q = 0; q < ns; ++q)
      for (int m = 0; m < static_cast<int>(cut->scnt[q]); ++m) {
        const size_t r = static_cast<size_t>(cut->smem[25 * static_cast<size_t>(q) + m]);
        for (int v = 0; v < 4; ++v) acc[v * g.plane + r] += Qh[static_cast<size_t>(q) * 4 + v];
      }
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const size_t k = static_cast<size_t>(j + NG) * g.mx + (i + NG);
        if (cut->vol[k] <= 0.0 || (cut->smallpos[k] < 0.0 && cut->nbrcount[k] == 1.0)) continue;
        for (int v = 0; v < 4; ++v) Uc[v * g.plane + k] = acc[v * g.plane + k] / cut->nbrcount[k];
      }
  }
}

// out = a A + b (B + dt R) on interior cells, dt per case.
void combine(const ugrid_setup& s, double a, const double* A, double b, const double* B, const double* R,
             const std::vector<double>& dt, double* out) {
  const Shape g(s);
  for (int c = 0; c < g.ncase; ++c)
    for (int v = 0; v < 4; ++v)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const size_t k = g.at(c, v, i, j);
          out[k] = a * A[k] + b * (B[k] + dt[c] * R[k]);
        }
}

int run(const ugrid_setup& s, double* U, double* t, double* dt_out, double tend, int64_t maxsteps, int64_t* nsteps,
        int64_t* nreject) {
  const Shape g(s);
  const size_t n = static_cast<size_t>(g.ncase) * 4 * g.plane;
  std::vector<double> mco(g.ncase, s.params.ausm_mco);
  if (s.mco)
    for (int c = 0; c < g.ncase; ++c) mco[c] = s.mco[c];
  // The reference's device arrays U1 / U2 / R start uninitialised; their ghosts are always
  // filled before use and only interiors are compared.
  std::vector<double> U1(U, U + n), U2(U, U + n), R(n, 0.0), dt(g.ncase);
  const ugrid_params& par = s.params;
  for (int64_t step = 0; step < maxsteps; ++step) {
    bool done = true;
    for (int c = 0; c < g.ncase; ++c)
      if (t[c] < tend) done = false;
    if (done) return 0;
    apply_bc(s, U);
    for (int c = 0; c < g.ncase; ++c) {
      double d = par.dt_fixed > 0.0 ? par.dt_fixed : par.cfl / bound_max(s, U, c);
      const double remain = tend - t[c];
      if (remain <= 0.0)
        d = 0.0;
      else if (d > remain)
        d = remain;
      dt[c] = d;
    }
    int retries = 0;
    for (;;) {
      residual(s, mco, U, R.data());
      combine(s, 0.0, U, 1.0, U, R.data(), dt, U1.data());
      bool bad = any_bad(s, U1.data());
      if (!bad && s.cut) {
        srd(s, U1.data());
        bad = any_bad(s, U1.data());
      }
      if (!bad) {
        apply_bc(s, U1.data());
        residual(s, mco, U1.data(), R.data());
        combine(s, 0.75, U, 0.25, U1.data(), R.data(), dt, U2.data());
        bad = any_bad(s, U2.data());
        if (!bad && s.cut) {
          srd(s, U2.data());
          bad = any_bad(s, U2.data());
        }
      }
      if (!bad) {
        apply_bc(s, U2.data());
        residual(s, mco, U2.data(), R.data());
        combine(s, 1.0 / 3.0, U, 2.0 / 3.0, U2.data(), R.data(), dt, U1.data());
        bad = any_bad(s, U1.data());
        if (!bad && s.cut) {
          srd(s, U1.data());
          bad = any_bad(s, U1.data());
        }
      }
      if (!bad) break;
      ++*nreject;
      ++retries;
      bool usable = retries <= 12;
      for (int c = 0; c < g.ncase; ++c) {
        if (dt[c] > 0.0) dt[c] *= 0.5;
        if (dt[c] < 0.0 || !std::isfinite(dt[c])) usable = false;
      }
      if (!usable) return 2;
    }
    for (int c = 0; c < g.ncase; ++c)
      for (int v = 0; v < 4; ++v)
        for (int j = 0; j < g.ny; ++j)
          for (int i = 0; i < g.nx; ++i) U[g.at(c, v, i, j)] = U1[g.at(c, v, i, j)];
    for (int c = 0; c < g.ncase; ++c) {
      dt_out[c] = dt[c];
      if (dt[c] > 0.0) t[c] += dt[c];
    }
    ++*nsteps;
  }
  return 1;
}

// ---------------------------------------------------------------- backward Euler --
struct Implicit {
  int64_t updates = 0, retries = 0, checks = 0;
  double initial_defect = 0, final_defect = 0, tolerance = 0;
  int reason = 0;  // 1 initial state, 2 pseudo retries, 3 non-finite defect, 4 not converged, 5 dt
};

bool implicit_controls_valid(const ugrid_setup& s, const double* t, const std::vector<double>& mco, double tend,
                             int64_t maxsteps) {
  const ugrid_params& p = s.params;
  if (!std::isfinite(tend) || maxsteps < 0 || !(p.gamma > 1.0) || !std::isfinite(p.gamma) || !(p.cfl > 0.0) ||
      !std::isfinite(p.cfl) || !(p.dt_fixed >= 0.0) || !std::isfinite(p.dt_fixed) || !(p.pseudo_cfl > 0.0) ||
      !std::isfinite(p.pseudo_cfl) || p.implicit_maxiter < 1 || !(p.implicit_rtol >= 0.0) ||
      !std::isfinite(p.implicit_rtol) || !(p.implicit_atol > 0.0) || !std::isfinite(p.implicit_atol))
    return false;
  for (int c = 0; c < s.ncase; ++c)
    if (!std::isfinite(t[c]) || !(mco[c] > 0.0) || !std::isfinite(mco[c])) return false;
  return true;
}

// Largest bound of case c on the bit patterns (the spectral rate of a regular cell is the
// wave-speed bound of the explicit step).
double rate_max(const ugrid_setup& s, const double* U, int c) { return bound_max(s, U, c); }

int run_implicit(const ugrid_setup& s, double* U, double* t, double* dt, double tend, int64_t maxsteps,
                 int64_t* nsteps, Implicit* st) {
  const Shape g(s);
  const ugrid_params& par = s.params;
  const double gm = par.gamma;
  std::vector<double> mco(g.ncase, par.ausm_mco);
  if (s.mco)
    for (int c = 0; c < g.ncase; ++c) mco[c] = s.mco[c];
  if (!implicit_controls_valid(s, t, mco, tend, maxsteps)) return 3;
  if (any_bad(s, U)) {
    st->reason = 1;
    return 2;
  }
  const size_t n = static_cast<size_t>(g.ncase) * 4 * g.plane;
  std::vector<double> Qa(n), Tb(n), R(n, 0.0), norms(g.ncase), initial(g.ncase);
  for (int64_t step = 0;; ++step) {
    bool done = true;
    for (int c = 0; c < g.ncase; ++c)
      if (t[c] < tend) done = false;
    if (done) return 0;
    if (step >= maxsteps) return 1;
    for (int c = 0; c < g.ncase; ++c) {
      const double rate = rate_max(s, U, c);
      const double remain = tend - t[c];
      if (remain <= 0.0) {
        dt[c] = 0.0;
        continue;
      }
      double d = par.dt_fixed > 0.0 ? par.dt_fixed : par.cfl / rate;
      if (!(d > 0.0) || !std::isfinite(d)) {
        st->reason = 5;
        return 3;
      }
      const double roundoff = 8 * std::numeric_limits<double>::epsilon() * std::fmax(std::fabs(tend), std::fabs(t[c]));
      if (d >= remain || remain - d <= roundoff) d = remain;
      dt[c] = d;
      if (!(t[c] + d > t[c])) {
        st->reason = 5;
        return 3;
      }
    }
    std::copy(U, U + n, Qa.begin());
    std::copy(U, U + n, Tb.begin());
    double* Q = Qa.data();
    double* trial = Tb.data();
    double pseudo = par.pseudo_cfl;
    for (int iter = 0;; ++iter) {
      apply_bc(s, Q);
      residual(s, mco, Q, R.data());
      for (int c = 0; c < g.ncase; ++c) {
        uint64_t best = 0;
        if (dt[c] > 0.0) {
          for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i) {
              if (s.cut && s.cut->vol[static_cast<size_t>(j + NG) * g.mx + (i + NG)] <= 0.0) continue;
              double norm = 0.0;
              for (int v = 0; v < 4; ++v) {
                const size_t k = g.at(c, v, i, j);
                const double defect = dt[c] * R[k] - (Q[k] - U[k]);
                if (!std::isfinite(defect)) {
                  norm = std::numeric_limits<double>::infinity();
                  break;
                }
                norm = std::fmax(norm, std::fabs(defect) / std::fmax(1.0, std::fabs(U[k])));
              }
              uint64_t bits;
              std::memcpy(&bits, &norm, 8);
              if (bits > best) best = bits;
            }
        }
        std::memcpy(&norms[c], &best, 8);
      }
      for (int c = 0; c < g.ncase; ++c)
        if (!std::isfinite(norms[c])) {
          st->reason = 3;
          return 2;
        }
      // implicit2d_converged
      bool converged = true;
      st->initial_defect = st->final_defect = st->tolerance = 0.0;
      for (int c = 0; c < g.ncase; ++c) {
        if (iter == 0) initial[c] = norms[c];
        const double tol = std::max(par.implicit_atol, par.implicit_rtol * initial[c]);
        st->initial_defect = std::max(st->initial_defect, initial[c]);
        st->final_defect = std::max(st->final_defect, norms[c]);
        st->tolerance = std::max(st->tolerance, tol);
        if (!std::isfinite(norms[c]) || !std::isfinite(tol) || norms[c] > tol) converged = false;
      }
      ++st->checks;
      if (converged) break;
      if (iter >= par.implicit_maxiter) {
        st->reason = 4;
        return 4;
      }
      int retries = 0;
      for (;;) {
        for (int c = 0; c < g.ncase; ++c) {
          if (dt[c] <= 0.0) continue;
          for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i) {
              const size_t kc = static_cast<size_t>(j + NG) * g.mx + (i + NG);
              if (s.cut && s.cut->vol[kc] <= 0.0) continue;
              const size_t k = g.at(c, 0, i, j);
              const double r = Q[k], mu = Q[k + g.plane], mv = Q[k + 2 * g.plane], E = Q[k + 3 * g.plane];
              const double u = mu / r, v = mv / r;
              const double p = (gm - 1.0) * (E - 0.5 * (mu * mu + mv * mv) / r);
              const double a = std::sqrt(gm * p / r);
              double rate;
              if (!s.cut) {
                rate = (std::fabs(u) + a) / s.dx[i + NG] + (std::fabs(v) + a) / s.dy[j + NG];
              } else {
                // implicit_spectral_rate of a cell of a cut grid (src/implicit2d.h:33-43)
                const ugrid_cut_view* ct = s.cut;
                const size_t fe = static_cast<size_t>(j) * (g.nx + 1) + i, fn = static_cast<size_t>(j) * g.nx + i;
                const double wall = std::sqrt(ct->wax[kc] * ct->wax[kc] + ct->way[kc] * ct->way[kc]);
                rate = ((std::fabs(u) + a) * (ct->axl[fe + 1] + ct->axl[fe]) +
                        (std::fabs(v) + a) * (ct->ayl[fn + g.nx] + ct->ayl[fn]) +
                        std::fabs(u * ct->wax[kc] + v * ct->way[kc]) + a * wall) /
                       (2.0 * ct->vol[kc]);
              }
              const double omega = pseudo / (pseudo + dt[c] * rate);
              for (int w = 0; w < 4; ++w) {
                const size_t q = k + w * g.plane;
                trial[q] = Q[q] + omega * (dt[c] * R[q] - (Q[q] - U[q]));
              }
            }
        }
        if (!any_bad(s, trial)) break;
        ++st->retries;
        if (++retries > 12) {
          st->reason = 2;
          return 2;
        }
        pseudo *= 0.5;
      }
      std::swap(Q, trial);
      ++st->updates;
    }
    for (int c = 0; c < g.ncase; ++c)
      for (int v = 0; v < 4; ++v)
        for (int j = 0; j < g.ny; ++j)
          for (int i = 0; i < g.nx; ++i) U[g.at(c, v, i, j)] = Q[g.at(c, v, i, j)];
    for (int c = 0; c < g.ncase; ++c) {
      if (dt[c] <= 0.0) continue;
      t[c] = dt[c] == tend - t[c] ? tend : t[c] + dt[c];
    }
    ++*nsteps;
  }
}

}  // namespace grid_oracle

extern "C" {

// Backward Euler as the reference's GPU backend runs it.  stats: updates, retries, checks
// (int64), then initial_defect, final_defect, tolerance (doubles), reason (int64); all in/out,
// counters cumulative.  Returns 0, 1, 2, 3 or 4.
int grid_oracle_run_implicit(const ugrid_setup* s, double* u, double* t, double* dt, double tend, int64_t maxsteps,
                             int64_t* nsteps, int64_t* counts, double* defects) {
  grid_oracle::Implicit st;
  st.updates = counts[0];
  st.retries = counts[1];
  st.checks = counts[2];
  st.initial_defect = defects[0];
  st.final_defect = defects[1];
  st.tolerance = defects[2];
  const int status = grid_oracle::run_implicit(*s, u, t, dt, tend, maxsteps, nsteps, &st);
  counts[0] = st.updates;
  counts[1] = st.retries;
  counts[2] = st.checks;
  counts[3] = st.reason;
  defects[0] = st.initial_defect;
  defects[1] = st.final_defect;
  defects[2] = st.tolerance;
  return status;
}


void grid_oracle_apply_bc(const ugrid_setup* s, double* u) { grid_oracle::apply_bc(*s, u); }

// Residual at u: boundary ghosts filled on a copy first (u is not modified).
int grid_oracle_rhs(const ugrid_setup* s, const double* u, double* r) {
  const grid_oracle::Shape g(*s);
  const size_t n = static_cast<size_t>(g.ncase) * 4 * g.plane;
  std::vector<double> U(u, u + n);
  grid_oracle::apply_bc(*s, U.data());
  std::vector<double> mco(g.ncase, s->params.ausm_mco);
  if (s->mco)
    for (int c = 0; c < g.ncase; ++c) mco[c] = s->mco[c];
  std::memset(r, 0, n * sizeof(double));
  grid_oracle::residual(*s, mco, U.data(), r);
  return 0;
}

// Advance u (ghosts included, as the reference's device array ends), t and dt (last accepted
// steps) to tend; nsteps / nreject are incremented.  Returns 0, 1 or 2 as the backend does.
int grid_oracle_run(const ugrid_setup* s, double* u, double* t, double* dt, double tend, int64_t maxsteps,
                    int64_t* nsteps, int64_t* nreject) {
  return grid_oracle::run(*s, u, t, dt, tend, maxsteps, nsteps, nreject);
}

// Legacy face flux of (rho, normal velocity, tangential velocity, p) states: kind 0 HLLC,
// 1 AUSM+-up.  F = mass, normal momentum, tangential momentum, energy.
void grid_oracle_flux(int kind, const double* L, const double* R, double gamma, double mco, double* F) {
  if (kind == 1)
    grid_oracle::ausm(L[0], L[1], L[2], L[3], R[0], R[1], R[2], R[3], gamma, mco, F);
  else
    grid_oracle::hllc(L[0], L[1], L[2], L[3], R[0], R[1], R[2], R[3], gamma, F);
}

}  // extern "C"
