// ugrid.hip -- batched first-order Euler solver on a shared tensor-product grid (gfx950, FP64):
// the device side of ugrid_hip.h, replacing the reference's legacy CUDA backend
// (src/euler2d_cuda.cu:153-447 kernels, :530-708 driver).
//
// Design.  The reference launches, per RK stage, separate primitive, x-flux, y-flux, collect,
// RK-combine and admissibility kernels with a host round trip per stage (flag copy) and per
// step (smax copy, dt upload).  Here one kernel per RK stage does all of it and the step
// control lives on the device:
//   * a wave owns a strip of 62 columns x `rows` rows of one case.  Lane k holds column
//     i0 - 1 + k, so lanes 0 and 63 are the strip's x halo; each lane computes the flux of the
//     face west of its cell (left state from lane k - 1 by a cross-lane shuffle) and takes its
//     east flux from lane k + 1, so every x face of the strip is evaluated once by 64 lanes at
//     a time.  The wave marches north through its rows, keeping the previous row's primitives,
//     x-residual and y-face flux in registers, so each y face is evaluated once and a cell's
//     state is read once per stage (plus one halo row per strip).  No LDS, no barriers.
//   * boundary ghosts are never materialised: a halo lane or row outside the domain reads the
//     mapped interior cell (periodic wrap, transmissive clamp, reflect mirror) and negates the
//     normal velocity for a reflection -- bit for bit the primitives of the reference's filled
//     ghost (negation is exact).
//   * the RK update, the admissibility test of the candidate and, in the last stage, the
//     candidate's wave-speed bound for the next step's dt are fused into the same pass.  A
//     one-wave kernel after the last stage commits or rejects the trial for all cases: t += dt, step counters, buffer swap (U[cur] and the other buffer alternate, so no
//     accept copy), next dt from the bound, or dt halving and the 12-retry limit.
//   * a trial is four launches; the host replays a graph of several trials and polls the
//     control record between replays.
// The arithmetic of every formula follows the reference's GPU kernels operation for operation
// and this file is compiled without FMA contraction, so results are bitwise those of the CPU
// restatement of that backend.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ugrid_hip.h"

namespace ugrid {

constexpr int kLanes = 64;
constexpr int kStripCells = kLanes - 2;  // columns a strip updates
constexpr int kWaves = 4;                // strips per block
constexpr int kThreads = kLanes * kWaves;
constexpr int kGhosts = 2;
constexpr int kXcds = 8;
constexpr int kMaxRetries = 12;
// First-stage pre-checks per trial on a cut grid (the body bench needs about six halvings a step).
constexpr int kPrechecks = 8;
enum : int { BC_PERIODIC = 0, BC_TRANSMISSIVE = 1, BC_REFLECT = 2 };
enum : int { RUNNING = -1 };
// Backward-Euler phases of a physical step (each kernel of the replayed sequence acts only in its
// phase; the one-wave decision kernels move the phase).
enum : int { PH_BOUND = 0, PH_COPY = 1, PH_RES = 2, PH_TRIAL = 3 };
// Why a backward-Euler run stopped with 2, 3 or 4.
enum : int { WHY_NONE = 0, WHY_INITIAL = 1, WHY_PSEUDO = 2, WHY_DEFECT = 3, WHY_MAXITER = 4, WHY_DT = 5 };

// Step control of a batch (one record, device memory).
struct Control {
  double tend;
  long long maxsteps;  // steps allowed in the current run call
  long long steps;     // accepted in the current run call
  long long nsteps, nreject;
  long long events;    // decisions taken (every commit, rejection, phase change): progress
  int status;      // RUNNING, or the run's result 0 / 1 / 2 / 3 / 4 (also 0 before the first run)
  int bad;         // a stage found a nonphysical candidate cell (this trial)
  int retries;     // rejections of the current step
  int cur;         // B[cur] holds the accepted state
  int ghost_from;  // download ghosts: 0 as uploaded, 1 boundary fill of B[cur], 2 of B[cur ^ 1]
  // backward Euler
  int phase, iter, pretries, qbuf, reason;
  int pre_done;    // SSP-RK3 on a cut grid: this trial's first-stage pre-check has passed
  double pseudo;   // pseudo CFL of the current physical step
  long long updates, iretries, checks;
  double initial_defect, final_defect, tolerance;
};

struct Grid {
  int nx, ny, mx, my, ncase;
  int rows;            // rows per strip
  int strips_x, strips_y, strips;
  int blocks;          // blocks that own strips (the launch is padded to a multiple of 8)
  int bc[4];
  int ausm;
  double gamma, cfl, dt_fixed;
  size_t plane;        // mx * my
  const double* dx;    // mx widths (ghosts included)
  const double* dy;    // my
  const double* mco;   // ncase
  // State buffers: B[0] and B[1] alternate as accepted state and RK candidate, B[2] holds the
  // second RK stage; backward Euler rotates all three (accepted state, iterate, candidate).
  double* B[3];
  double* R;           // residual of the backward-Euler iterate
  double* t;
  double* dt;          // step of the running trial
  double* dt_out;      // last accepted step
  unsigned long long* smax;  // per case: bit pattern of the largest wave-speed bound (non-negative)
  unsigned long long* norm;  // per case: bit pattern of the largest scaled defect (backward Euler)
  double* initial;           // per case: first defect norm of the physical step
  double pseudo0, rtol, atol;
  int maxiter;
  // Cut cells (has_cut): areas, 0 / 1 / 2 flags, apertures, wall vectors, SRD counts and
  // positions; SRD neighbourhoods as member lists, and the cells SRD rewrites with the
  // neighbourhoods each belongs to (ascending), so the sums run in one fixed order.
  int has_cut, nsmall, n_srd;
  const double* vol;
  const double* flag;
  const double* axl;
  const double* ayl;
  const double* wax;
  const double* way;
  const double* nbrcount;
  const double* smallpos;
  const double* svhat;
  const int* mem_off;
  const int* mem_idx;
  const int* srd_cell;
  const int* grp_off;
  const int* grp_idx;
  double* Qh;  // ncase x nsmall x 4 neighbourhood averages
  // One byte per plane cell for the stage kernels (so that they read the double arrays above only
  // where a value is needed): bit 0 solid (vol <= 0), bit 1 cut (flag != 0), bit 2 rewritten by
  // SRD (smallpos >= 0 or nbrcount != 1), bits 3-4 the west x face, bits 5-6 the south y face of
  // the cell: 0 aperture +0, 1 aperture equal to the full width (dy, dx; the builders store that
  // value for open faces), 2 another value (read from axl / ayl).
  const uint8_t* code;
  // The interior cut cells (plane indices, ascending): the first RK stage is tried on them alone
  // before the whole-grid pass (k_stage1_precheck).
  const int* pre_cell;
  int n_pre;
  Control* ctrl;
};

struct Prim {
  double r, u, v, p;
};
struct Flux {
  double m, x, y, e;  // mass, x momentum, y momentum, energy
};

// ---------------------------------------------------------------------------------------------
// Pointwise formulas (operation order of src/euler2d_cuda.cu:71-146, 169-205, 248-281).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double pressure(double r, double mu, double mv, double E, double gamma) {
  return (gamma - 1.0) * (E - 0.5 * (mu * mu + mv * mv) / r);
}

__device__ __forceinline__ Prim primitive(double r, double mu, double mv, double E, double gamma) {
  return {r, mu / r, mv / r, pressure(r, mu, mv, E, gamma)};
}

// Exact flux along the normal velocity un (tangential ut): mass, normal, tangential, energy.
__device__ __forceinline__ void exact_flux(double r, double un, double ut, double p, double gm, double f[4]) {
  const double E = p / (gm - 1.0) + 0.5 * r * (un * un + ut * ut);
  f[0] = r * un;
  f[1] = r * un * un + p;
  f[2] = r * un * ut;
  f[3] = un * (E + p);
}

__device__ void hllc(double rL, double uL, double vL, double pL, double rR, double uR, double vR, double pR, double gm,
                     double f[4]) {
  const double aL = sqrt(gm * pL / rL), aR = sqrt(gm * pR / rR);
  const double EL = pL / (gm - 1.0) + 0.5 * rL * (uL * uL + vL * vL);
  const double ER = pR / (gm - 1.0) + 0.5 * rR * (uR * uR + vR * vR);
  const double qL = sqrt(rL), qR = sqrt(rR);
  const double HL = (EL + pL) / rL, HR = (ER + pR) / rR;
  const double uT = (qL * uL + qR * uR) / (qL + qR);
  const double vT = (qL * vL + qR * vR) / (qL + qR);
  const double HT = (qL * HL + qR * HR) / (qL + qR);
  const double aT = sqrt(fmax((gm - 1.0) * (HT - 0.5 * (uT * uT + vT * vT)), 1e-300));
  const double SL = fmin(uL - aL, uT - aT), SR = fmax(uR + aR, uT + aT);
  double fL[4], fR[4];
  exact_flux(rL, uL, vL, pL, gm, fL);
  exact_flux(rR, uR, vR, pR, gm, fR);
  if (SL >= 0.0) {
    for (int k = 0; k < 4; ++k) f[k] = fL[k];
    return;
  }
  if (SR <= 0.0) {
    for (int k = 0; k < 4; ++k) f[k] = fR[k];
    return;
  }
  const double Ss = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR));
  // Star state of the upwind side of the contact, then F = F_side + S_side (U* - U_side).
  const bool left = Ss >= 0.0;
  const double r = left ? rL : rR, u = left ? uL : uR, v = left ? vL : vR, p = left ? pL : pR;
  const double E = left ? EL : ER, S = left ? SL : SR;
  const double* fs = left ? fL : fR;
  const double a = r * (S - u) / (S - Ss);
  const double star[4] = {a, a * Ss, a * v, a * (E / r + (Ss - u) * (Ss + p / (r * (S - u))))};
  const double side[4] = {r, r * u, r * v, E};
  for (int k = 0; k < 4; ++k) f[k] = fs[k] + S * (star[k] - side[k]);
}

// AUSM+-up split functions (Liou 2006) with the reference's constants (beta 1/8, Kp 1/4,
// Ku 3/4, sigma 1).  s is the side sign +1 / -1.
__device__ __forceinline__ double m1(double m, int s) { return 0.5 * (m + s * fabs(m)); }
__device__ __forceinline__ double m2(double m, int s) { return s * 0.25 * (m + s) * (m + s); }
__device__ __forceinline__ double m4(double m, int s) {
  if (fabs(m) >= 1.0) return m1(m, s);
  return m2(m, s) * (1.0 - s * 2.0 * m2(m, -s));
}
__device__ __forceinline__ double p5(double m, int s, double alpha) {
  if (fabs(m) >= 1.0) return 0.5 * (1.0 + s * (m > 0.0 ? 1.0 : (m < 0.0 ? -1.0 : 0.0)));
  return m2(m, s) * ((s * 2.0 - m) - s * 16.0 * alpha * m * m2(m, -s));
}

__device__ void ausm(double rL, double uL, double vL, double pL, double rR, double uR, double vR, double pR, double gm,
                     double mco, double f[4]) {
  const double a12 = 0.5 * (sqrt(gm * pL / rL) + sqrt(gm * pR / rR));
  const double rho12 = 0.5 * (rL + rR), ML = uL / a12, MR = uR / a12;
  const double mb2 = 0.5 * (uL * uL + uR * uR) / (a12 * a12);
  const double m02 = fmin(1.0, fmax(mb2, mco * mco));
  const double fa = sqrt(m02) * (2.0 - sqrt(m02));
  const double alpha = 3.0 / 16.0 * (-4.0 + 5.0 * fa * fa);
  const double m12 =
      m4(ML, 1) + m4(MR, -1) - (0.25 / fa) * fmax(1.0 - mb2, 0.0) * (pR - pL) / (rho12 * a12 * a12);
  const double pp = p5(ML, 1, alpha), pm = p5(MR, -1, alpha);
  const double p12 = pp * pL + pm * pR - 0.75 * pp * pm * (rL + rR) * fa * a12 * (uR - uL);
  const bool left = m12 > 0.0;
  const double r = left ? rL : rR, u = left ? uL : uR;
  const double v = left ? vL : vR, p = left ? pL : pR;
  const double mdot = a12 * m12 * r;
  const double E = p / (gm - 1.0) + 0.5 * r * (u * u + v * v);
  f[0] = mdot;
  f[1] = mdot * u + p12;
  f[2] = mdot * v;
  f[3] = mdot * (E + p) / r;
}

// Flux through an x face (normal u) or a y face (normal v), returned in x / y momentum order.
template <bool XFACE>
__device__ __forceinline__ Flux face_flux(const Grid& g, double mco, const Prim& L, const Prim& R) {
  double f[4];
  const double nL = XFACE ? L.u : L.v, tL = XFACE ? L.v : L.u;
  const double nR = XFACE ? R.u : R.v, tR = XFACE ? R.v : R.u;
  if (g.ausm)
    ausm(L.r, nL, tL, L.p, R.r, nR, tR, R.p, g.gamma, mco, f);
  else
    hllc(L.r, nL, tL, L.p, R.r, nR, tR, R.p, g.gamma, f);
  return XFACE ? Flux{f[0], f[1], f[2], f[3]} : Flux{f[0], f[2], f[1], f[3]};
}

__device__ __forceinline__ double wave_bound(double r, double mu, double mv, double E, const Grid& g, double dx,
                                             double dy) {
  const double u = mu / r, v = mv / r;
  const double p = pressure(r, mu, mv, E, g.gamma);
  const double a = sqrt(g.gamma * p / r);
  return (fabs(u) + a) / dx + (fabs(v) + a) / dy;
}

__device__ __forceinline__ bool inadmissible(double r, double mu, double mv, double E, double gamma) {
  const double p = pressure(r, mu, mv, E, gamma);
  return !(r > 0.0) || !(p > 0.0) || !isfinite(r) || !isfinite(p);
}

// Ghost index map of one axis (src/euler2d_cuda.cu:153-180): the interior cell a ghost copies
// and whether its normal momentum is negated.  Indices beyond the first ghost (unused lanes)
// are clamped.
__device__ __forceinline__ int map_axis(int i, int n, int lo, int hi, bool* flip) {
  if (i < 0) {
    if (lo == BC_PERIODIC) return i + n;
    if (lo == BC_REFLECT) {
      *flip = true;
      return -i - 1;
    }
    return 0;
  }
  if (i >= n) {
    if (i > n) return n - 1;
    if (hi == BC_PERIODIC) return i - n;
    if (hi == BC_REFLECT) {
      *flip = true;
      return 2 * n - 1 - i;
    }
    return n - 1;
  }
  return i;
}

__device__ __forceinline__ double shfl_up1(double x) { return __shfl_up(x, 1, kLanes); }
__device__ __forceinline__ double shfl_down1(double x) { return __shfl_down(x, 1, kLanes); }

__device__ __forceinline__ unsigned long long wave_max(unsigned long long x) {
  for (int d = kLanes / 2; d > 0; d >>= 1) {
    const unsigned long long y = __shfl_xor(x, d, kLanes);
    x = y > x ? y : x;
  }
  return x;
}

__device__ __forceinline__ double next_dt(const Grid& g, double smax_value, double t, double tend) {
  double d = g.dt_fixed > 0.0 ? g.dt_fixed : g.cfl / smax_value;
  const double remain = tend - t;
  if (remain <= 0.0)
    d = 0.0;
  else if (d > remain)
    d = remain;
  return d;
}

__device__ __forceinline__ unsigned long long load_relaxed(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Logical block of a launch: hardware dispatches consecutive blocks round-robin over the 8 XCDs;
// give each XCD a contiguous range of strips so that halo rows and columns shared by neighbours
// stay in one L2.  The launch is a multiple of 8 blocks.
__device__ __forceinline__ int logical_block() {
  const int per_xcd = gridDim.x / kXcds;
  return (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
}

// ---------------------------------------------------------------------------------------------
// Commit of a trial (one wave, after the last stage): src/euler2d_cuda.cu:670-693.
// ---------------------------------------------------------------------------------------------
// Rejected trial (one wave): halve every positive step; give up after 12 retries or on an
// unusable step.
__device__ void reject(const Grid& g) {
  Control* c = g.ctrl;
  const int lane = threadIdx.x % kLanes;
  int unusable = 0;
  for (int k = lane; k < g.ncase; k += kLanes) {
    double d = g.dt[k];
    if (d > 0.0) d *= 0.5;
    g.dt[k] = d;
    if (d < 0.0 || !isfinite(d)) unusable = 1;
    g.smax[k] = 0;
  }
  unusable = __any(unusable);
  if (lane == 0) {
    const int retries = c->retries + 1;
    c->retries = retries;
    c->nreject += 1;
    if (retries > kMaxRetries || unusable) c->status = 2;
    c->bad = 0;
    c->pre_done = 0;
    c->events += 1;
  }
}

__device__ void commit(const Grid& g) {
  Control* c = g.ctrl;
  const int lane = threadIdx.x % kLanes;
  if (__hip_atomic_load(&c->bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
    reject(g);
    return;
  }
  // Accepted: advance the clocks of the cases that moved, then either stop (maxsteps reached,
  // or every case at tend) or choose the next step from the candidate's wave-speed bound.
  const double tend = c->tend;
  const long long steps = c->steps + 1;
  const bool maxed = steps >= c->maxsteps;
  int open = 0;
  for (int k = lane; k < g.ncase; k += kLanes) {
    const double d = g.dt[k];
    g.dt_out[k] = d;
    double tk = g.t[k];
    if (d > 0.0) {
      tk += d;
      g.t[k] = tk;
    }
    if (tk < tend) open = 1;
  }
  const bool next = !maxed && __any(open);
  for (int k = lane; k < g.ncase; k += kLanes) {
    if (next) g.dt[k] = next_dt(g, __longlong_as_double(static_cast<long long>(load_relaxed(&g.smax[k]))), g.t[k], tend);
    g.smax[k] = 0;
  }
  if (lane == 0) {
    c->steps = steps;
    c->cur ^= 1;
    c->nsteps += 1;
    c->retries = 0;
    c->pre_done = 0;
    c->ghost_from = next ? 1 : 2;
    if (!next) c->status = maxed ? 1 : 0;
    c->events += 1;
  }
}

// ---------------------------------------------------------------------------------------------
// One RK stage (MODE 1..3), the residual alone (MODE 0) or the backward-Euler residual with the
// defect norms (MODE 4), fused over a wave strip.
//   MODE 1: U[cur^1] = U + dt L(U)                    (in U[cur])
//   MODE 2: S2 = 3/4 U + 1/4 (U1 + dt L(U1))          (U1 in U[cur^1])
//   MODE 3: U[cur^1] = 1/3 U + 2/3 (S2 + dt L(S2)), and the candidate's wave-speed bound
//   MODE 0: R = L(in)
//   MODE 4: R = L(Q), per case max over cells of max_v |dt R - (Q - U)| / max(1, |U|)
//           (infinite for a non-finite defect), src/euler2d_cuda.cu:331-350
// ---------------------------------------------------------------------------------------------
// The cut-grid instantiation takes 140-144 VGPRs (3 waves per SIMD against the regular kernels'
// 4); holding it to 4 waves spills 10-13 registers and was measured slower (2.40 against 2.37 ms
// per step on 16 x 1024^2 with a body).
template <int MODE, bool CUT>
__global__ __launch_bounds__(kThreads) void k_stage(Grid g, const double* __restrict__ rhs_in,
                                                     double* __restrict__ rhs_out) {
  Control* ctrl = g.ctrl;
  bool skip = false;
  int cur = 0, qb = 0;
  if (MODE > 0) {
    if (__hip_atomic_load(&ctrl->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != RUNNING) return;
    if (MODE == 4) {
      const int phase = ctrl->phase;
      if (phase != PH_COPY && phase != PH_RES) return;
      qb = ctrl->qbuf;
    } else {
      skip = __hip_atomic_load(&ctrl->bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    }
    cur = ctrl->cur;
  }
  const double* base = MODE > 0 ? g.B[cur] : nullptr;
  const double* in = MODE == 0 ? rhs_in : MODE == 1 ? g.B[cur] : MODE == 2 ? g.B[cur ^ 1] : MODE == 3 ? g.B[2]
                                                                                                    : g.B[qb];
  double* out = MODE == 0 ? rhs_out : MODE == 2 ? g.B[2] : MODE == 4 ? g.R : g.B[cur ^ 1];
  constexpr double A = MODE == 2 ? 0.75 : MODE == 3 ? 1.0 / 3.0 : 0.0;
  constexpr double B = MODE == 2 ? 0.25 : MODE == 3 ? 2.0 / 3.0 : 1.0;

  const int lane = threadIdx.x % kLanes;
  const int strip = logical_block() * kWaves + threadIdx.x / kLanes;
  bool wave_bad = false;
  unsigned long long bound = 0;
  const bool work = !skip && strip < g.strips;
  int icase = 0;
  if (work) {
    icase = strip / (g.strips_x * g.strips_y);
    const int rem = strip - icase * g.strips_x * g.strips_y;
    const int sy = rem / g.strips_x, sx = rem - sy * g.strips_x;
    const int i = sx * kStripCells - 1 + lane;
    const int j0 = sy * g.rows;
    const int jl = min(j0 + g.rows, g.ny) - 1;
    const bool owner = lane >= 1 && lane <= kStripCells && i < g.nx;
    bool flip_u = false;
    const int ic = map_axis(i, g.nx, g.bc[0], g.bc[1], &flip_u);
    const size_t plane = g.plane;
    const size_t case_base = static_cast<size_t>(icase) * 4 * plane + (ic + kGhosts);
    const double mco = g.mco[icase];
    const double dxi = g.dx[min(max(i, -kGhosts), g.nx + kGhosts - 1) + kGhosts];
    const double dt = MODE > 0 ? g.dt[icase] : 0.0;

    Prim wprev{};
    Flux gprev{}, xres{};  // xres: x residual of a full cell, or ae FX_e - aw FX_w of a cut cell
    double ay_prev = 1.0;  // aperture of the previous y face (cut grids)
    double q0 = 0, q1 = 0, q2 = 0, q3 = 0;  // conserved state of the owned cell of the previous row
    unsigned code_prev = 0;                  // cut grids: code of the lane's cell of the previous row
    const bool in_x = i >= 0 && i <= g.nx;   // the lane's west x face is a grid face
    for (int j = j0 - 1; j <= jl + 1; ++j) {
      bool flip_v = false;
      const int jc = map_axis(j, g.ny, g.bc[2], g.bc[3], &flip_v);
      const size_t k = case_base + static_cast<size_t>(jc + kGhosts) * g.mx;
      const double c0 = in[k], c1 = in[k + plane], c2 = in[k + 2 * plane], c3 = in[k + 3 * plane];
      // Cut grids: codes of the mapped cell (solid or not) and of the lane's own column at row j
      // (its faces, cut or not, rewritten by SRD or not).
      const unsigned code_m = CUT ? g.code[static_cast<size_t>(jc + kGhosts) * g.mx + ic + kGhosts] : 0u;
      const unsigned code_u = CUT && in_x ? g.code[static_cast<size_t>(j + kGhosts) * g.mx + i + kGhosts] : 0u;
      Prim w;
      if (CUT && (code_m & 1u)) {
        w = Prim{1.0, 0.0, 0.0, 1.0};  // solid cell (never a ghost's source: the body keeps clear of the sides)
      } else {
        w = primitive(c0, c1, c2, c3, g.gamma);
        if (flip_u) w.u = -w.u;
        if (flip_v) w.v = -w.v;
      }
      if (j >= j0) {
        // y face between rows j - 1 and j (closed faces carry no flux)
        double ay = 1.0;
        if (CUT && i >= 0 && i < g.nx) {
          const unsigned st = (code_u >> 5) & 3u;
          ay = st == 0u ? 0.0 : st == 1u ? g.dx[i + kGhosts] : g.ayl[static_cast<size_t>(j) * g.nx + i];
        }
        Flux gf{0.0, 0.0, 0.0, 0.0};
        if (!CUT || ay > 0.0) gf = face_flux<false>(g, mco, wprev, w);
        if (j > j0 && owner) {
          // Residual and update of the owned cell of row j - 1.
          const double dyj = g.dy[j - 1 + kGhosts];
          const size_t kc = static_cast<size_t>(j - 1 + kGhosts) * g.mx + i + kGhosts;
          double r0, r1, r2, r3;
          bool fluid = true, moved = false;
          if (CUT && (code_prev & 1u)) {
            fluid = false;
            r0 = r1 = r2 = r3 = 0.0;
          } else if (!CUT || !(code_prev & 2u)) {
            r0 = xres.m - (gf.m - gprev.m) / dyj;
            r1 = xres.x - (gf.x - gprev.x) / dyj;
            r2 = xres.y - (gf.y - gprev.y) / dyj;
            r3 = xres.e - (gf.e - gprev.e) / dyj;
          } else {
            const double iv = 1.0 / g.vol[kc], pw = wprev.p;
            r0 = -iv * (xres.m + ay * gf.m - ay_prev * gprev.m);
            r1 = -iv * (xres.x + ay * gf.x - ay_prev * gprev.x);
            r2 = -iv * (xres.y + ay * gf.y - ay_prev * gprev.y);
            r3 = -iv * (xres.e + ay * gf.e - ay_prev * gprev.e);
            r1 -= iv * pw * g.wax[kc];
            r2 -= iv * pw * g.way[kc];
          }
          if (CUT && MODE == 3) moved = (code_prev & 4u) != 0u;
          const size_t ko = case_base + static_cast<size_t>(j - 1 + kGhosts) * g.mx;
          if (MODE == 0 || MODE == 4) {
            out[ko] = r0;
            out[ko + plane] = r1;
            out[ko + 2 * plane] = r2;
            out[ko + 3 * plane] = r3;
            if (MODE == 4 && dt > 0.0 && fluid) {
              const double r[4] = {r0, r1, r2, r3}, q[4] = {q0, q1, q2, q3};
              double norm = 0.0;
              for (int v = 0; v < 4; ++v) {
                const double old = base[ko + v * plane];
                const double defect = dt * r[v] - (q[v] - old);
                if (!isfinite(defect)) {
                  norm = HUGE_VAL;
                  break;
                }
                norm = fmax(norm, fabs(defect) / fmax(1.0, fabs(old)));
              }
              const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(norm));
              bound = bits > bound ? bits : bound;
            }
          } else {
            double a0 = q0, a1 = q1, a2 = q2, a3 = q3;
            if (MODE > 1) {
              a0 = __builtin_nontemporal_load(base + ko);
              a1 = __builtin_nontemporal_load(base + ko + plane);
              a2 = __builtin_nontemporal_load(base + ko + 2 * plane);
              a3 = __builtin_nontemporal_load(base + ko + 3 * plane);
            }
            const double n0 = A * a0 + B * (q0 + dt * r0);
            const double n1 = A * a1 + B * (q1 + dt * r1);
            const double n2 = A * a2 + B * (q2 + dt * r2);
            const double n3 = A * a3 + B * (q3 + dt * r3);
            out[ko] = n0;
            out[ko + plane] = n1;
            out[ko + 2 * plane] = n2;
            out[ko + 3 * plane] = n3;
            if (fluid && inadmissible(n0, n1, n2, n3, g.gamma)) wave_bad = true;
            // the next step's bound (cells that SRD rewrites report theirs after it)
            if (MODE == 3 && fluid && !moved) {
              const double b = wave_bound(n0, n1, n2, n3, g, dxi, dyj);
              const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(b));
              bound = bits > bound ? bits : bound;
            }
          }
        }
        gprev = gf;
        ay_prev = ay;
      }
      if (j >= j0 && j <= jl) {
        // x faces of row j: lane k evaluates the face west of its cell.
        Prim wl;
        wl.r = shfl_up1(w.r);
        wl.u = shfl_up1(w.u);
        wl.v = shfl_up1(w.v);
        wl.p = shfl_up1(w.p);
        if (lane == 0) wl = w;
        double ax = 1.0;
        if (CUT && in_x) {
          const unsigned st = (code_u >> 3) & 3u;
          ax = st == 0u ? 0.0 : st == 1u ? g.dy[j + kGhosts] : g.axl[static_cast<size_t>(j) * (g.nx + 1) + i];
        }
        Flux fw{0.0, 0.0, 0.0, 0.0};
        if (!CUT || ax > 0.0) fw = face_flux<true>(g, mco, wl, w);
        const Flux fe{shfl_down1(fw.m), shfl_down1(fw.x), shfl_down1(fw.y), shfl_down1(fw.e)};
        const double ae = CUT ? shfl_down1(ax) : 1.0;
        if (CUT && owner && (code_u & 2u))
          xres = Flux{ae * fe.m - ax * fw.m, ae * fe.x - ax * fw.x, ae * fe.y - ax * fw.y, ae * fe.e - ax * fw.e};
        else
          xres = Flux{-(fe.m - fw.m) / dxi, -(fe.x - fw.x) / dxi, -(fe.y - fw.y) / dxi, -(fe.e - fw.e) / dxi};
      }
      wprev = w;
      q0 = c0, q1 = c1, q2 = c2, q3 = c3;
      code_prev = code_u;
    }
  }
  if (MODE == 0) return;
  if (MODE == 4) {
    bound = wave_max(bound);
    if (work && lane == 0 && bound) atomicMax(&g.norm[icase], bound);
    return;
  }
  if (__any(wave_bad) && lane == 0) atomicOr(&ctrl->bad, 1);
  if (MODE == 3) {
    bound = wave_max(bound);
    if (work && lane == 0 && bound) atomicMax(&g.smax[icase], bound);
  }
}

// The commit of a trial, one wave after the last stage (the kernel boundary orders every
// stage's flags and bounds before it; an in-kernel last-arriver commit would need device-scope
// fences, which write back and invalidate an XCD's L2 and cost more than this launch).
__global__ __launch_bounds__(kLanes) void k_commit(Grid g) {
  if (g.ctrl->status != RUNNING) return;
  commit(g);
}

// Wave-speed bound of every case at U[cur] (the first step of a run, src/euler2d_cuda.cu:223-237).
// Block b serves case b / per_case; each block reduces its cells to one maximum (on the bit
// patterns, as the reference's atomicMax does) and issues one atomic.
// gate >= 0: only while a run is in that backward-Euler phase.
__global__ __launch_bounds__(kThreads) void k_bound(Grid g, int per_case, int gate) {
  if (gate >= 0 && (g.ctrl->status != RUNNING || g.ctrl->phase != gate)) return;
  const int icase = blockIdx.x / per_case;
  const int part = blockIdx.x - icase * per_case;
  const double* u = g.B[g.ctrl->cur] + static_cast<size_t>(icase) * 4 * g.plane;
  const size_t inner = static_cast<size_t>(g.nx) * g.ny;
  unsigned long long best = 0;
  for (size_t q = part * static_cast<size_t>(blockDim.x) + threadIdx.x; q < inner;
       q += static_cast<size_t>(per_case) * blockDim.x) {
    const int j = static_cast<int>(q / g.nx), i = static_cast<int>(q - static_cast<size_t>(j) * g.nx);
    const size_t k = static_cast<size_t>(j + kGhosts) * g.mx + i + kGhosts;
    if (g.has_cut && g.vol[k] <= 0.0) continue;
    const double b = wave_bound(u[k], u[k + g.plane], u[k + 2 * g.plane], u[k + 3 * g.plane], g, g.dx[i + kGhosts],
                                g.dy[j + kGhosts]);
    const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(b));
    best = bits > best ? bits : best;
  }
  best = wave_max(best);
  __shared__ unsigned long long s_best[kWaves];
  if (threadIdx.x % kLanes == 0) s_best[threadIdx.x / kLanes] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kWaves; ++w) best = s_best[w] > best ? s_best[w] : best;
    if (best) atomicMax(&g.smax[icase], best);
  }
}

// Start of a run call (src/euler2d_cuda.cu:605-619 for the first step): done / maxsteps tests
// and the first step's dt of every case.  One block.
__global__ __launch_bounds__(kThreads) void k_open(Grid g, double tend, long long maxsteps) {
  Control* c = g.ctrl;
  __shared__ int s_open;
  const int tid = threadIdx.x;
  if (tid == 0) s_open = 0;
  __syncthreads();
  for (int k = tid; k < g.ncase; k += blockDim.x)
    if (g.t[k] < tend) s_open = 1;
  __syncthreads();
  const bool run = maxsteps > 0 && s_open;
  for (int k = tid; k < g.ncase; k += blockDim.x) {
    if (run) g.dt[k] = next_dt(g, __longlong_as_double(static_cast<long long>(g.smax[k])), g.t[k], tend);
    g.smax[k] = 0;
  }
  __syncthreads();
  if (tid == 0) {
    c->tend = tend;
    c->maxsteps = maxsteps;
    c->steps = 0;
    c->retries = 0;
    c->bad = 0;
    c->pre_done = 0;
    if (maxsteps <= 0)
      c->status = 1;
    else if (!s_open)
      c->status = 0;
    else {
      c->status = RUNNING;
      c->ghost_from = 1;
    }
  }
}

// ------------------------------------------------------------------ backward Euler --
// src/euler2d_cuda.cu:370-447 with src/implicit2d.h:16-63 and src/implicit2d.cpp:37-69.  A replay
// runs k_bound, k_be_open, k_be_copy, k_stage<4>, k_be_decide, k_be_trial, k_be_check; each acts
// only in its phase, so one replay completes at least one phase.

// Nonphysical interior cells of B[cur] (the run's initial check).  Grid-stride over all cases.
__global__ __launch_bounds__(kThreads) void k_check_state(Grid g) {
  const double* u = g.B[g.ctrl->cur];
  const size_t inner = static_cast<size_t>(g.nx) * g.ny;
  bool bad = false;
  for (size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; q < inner * g.ncase;
       q += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t c = q / inner, ij = q - c * inner;
    const int j = static_cast<int>(ij / g.nx), i = static_cast<int>(ij - static_cast<size_t>(j) * g.nx);
    const size_t kc = static_cast<size_t>(j + kGhosts) * g.mx + i + kGhosts;
    if (g.has_cut && g.vol[kc] <= 0.0) continue;
    const size_t k = c * 4 * g.plane + kc;
    if (inadmissible(u[k], u[k + g.plane], u[k + 2 * g.plane], u[k + 3 * g.plane], g.gamma)) bad = true;
  }
  if (__any(bad) && threadIdx.x % kLanes == 0) atomicOr(&g.ctrl->bad, 1);
}

__global__ __launch_bounds__(kLanes) void k_be_start(Grid g, double tend, long long maxsteps) {
  Control* c = g.ctrl;
  if (threadIdx.x != 0) return;
  c->tend = tend;
  c->maxsteps = maxsteps;
  c->steps = 0;
  c->reason = WHY_NONE;
  if (c->bad) {
    c->status = 2;
    c->reason = WHY_INITIAL;
  } else {
    c->status = RUNNING;
    c->phase = PH_BOUND;
  }
  c->bad = 0;
  c->events += 1;
}

// Opening of a physical step: done / maxsteps tests, then every case's dt from its spectral
// bound (implicit2d_physical_dt: fixed or CFL step, cut to the remaining time when within
// 8 eps of it, and it must move the clock).  Cases are processed in order: those before the
// first failing one keep their new step.
__global__ __launch_bounds__(kLanes) void k_be_open(Grid g) {
  Control* c = g.ctrl;
  if (c->status != RUNNING || c->phase != PH_BOUND) return;
  const int lane = threadIdx.x;
  const double tend = c->tend;
  int open = 0;
  for (int k = lane; k < g.ncase; k += kLanes)
    if (g.t[k] < tend) open = 1;
  const bool any_open = __any(open);
  const bool maxed = c->steps >= c->maxsteps;
  int first_fail = g.ncase;
  if (any_open && !maxed) {
    for (int k = lane; k < g.ncase; k += kLanes) {
      const double t = g.t[k];
      const double remain = tend - t;
      if (remain <= 0.0) continue;
      double d = g.dt_fixed > 0.0 ? g.dt_fixed : g.cfl / __longlong_as_double(static_cast<long long>(g.smax[k]));
      if (!(d > 0.0) || !isfinite(d)) {
        first_fail = min(first_fail, k);
        continue;
      }
      const double roundoff = 8 * DBL_EPSILON * fmax(fabs(tend), fabs(t));
      if (d >= remain || remain - d <= roundoff) d = remain;
      if (!(t + d > t)) first_fail = min(first_fail, k);
    }
    for (int d = kLanes / 2; d > 0; d >>= 1) first_fail = min(first_fail, __shfl_xor(first_fail, d, kLanes));
    for (int k = lane; k < g.ncase; k += kLanes) {
      const double t = g.t[k];
      const double remain = tend - t;
      double d = 0.0;
      bool assign = true;
      if (remain > 0.0) {
        d = g.dt_fixed > 0.0 ? g.dt_fixed : g.cfl / __longlong_as_double(static_cast<long long>(g.smax[k]));
        if (!(d > 0.0) || !isfinite(d)) {
          assign = false;
        } else {
          const double roundoff = 8 * DBL_EPSILON * fmax(fabs(tend), fabs(t));
          if (d >= remain || remain - d <= roundoff) d = remain;
        }
      }
      if (assign && k <= first_fail) {
        g.dt[k] = d;
        g.dt_out[k] = d;
      }
    }
  }
  for (int k = lane; k < g.ncase; k += kLanes) g.smax[k] = 0;
  if (lane == 0) {
    if (!any_open) {
      c->status = 0;
    } else if (maxed) {
      c->status = 1;
    } else if (first_fail < g.ncase) {
      c->status = 3;
      c->reason = WHY_DT;
    } else {
      c->phase = PH_COPY;
      c->pseudo = g.pseudo0;
      c->iter = 0;
      c->pretries = 0;
      c->qbuf = c->cur == 0 ? 1 : 0;
    }
    c->events += 1;
  }
}

// Iterate and candidate start as copies of the accepted state (interior cells).
__global__ __launch_bounds__(kThreads) void k_be_copy(Grid g) {
  const Control* c = g.ctrl;
  if (c->status != RUNNING || c->phase != PH_COPY) return;
  const int cur = c->cur, qb = c->qbuf, tb = 3 - cur - qb;
  const double* u = g.B[cur];
  double* q = g.B[qb];
  double* t = g.B[tb];
  const size_t inner = static_cast<size_t>(g.nx) * g.ny;
  for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < inner * 4 * g.ncase;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t cv = e / inner, ij = e - cv * inner;
    const int j = static_cast<int>(ij / g.nx), i = static_cast<int>(ij - static_cast<size_t>(j) * g.nx);
    const size_t k = cv * g.plane + static_cast<size_t>(j + kGhosts) * g.mx + i + kGhosts;
    const double x = u[k];
    q[k] = x;
    t[k] = x;
  }
}

// Convergence of the pseudo iteration (implicit2d_converged): tolerance max(atol, rtol * the
// step's first norm) per case; on convergence the iterate becomes the accepted state and the
// clocks advance (implicit2d_advance_time), else a candidate follows or the iteration limit
// ends the run with 4.
__global__ __launch_bounds__(kLanes) void k_be_decide(Grid g) {
  Control* c = g.ctrl;
  if (c->status != RUNNING || (c->phase != PH_COPY && c->phase != PH_RES)) return;
  const int lane = threadIdx.x;
  int nonfinite = 0;
  for (int k = lane; k < g.ncase; k += kLanes)
    if (!isfinite(__longlong_as_double(static_cast<long long>(g.norm[k])))) nonfinite = 1;
  if (__any(nonfinite)) {
    for (int k = lane; k < g.ncase; k += kLanes) g.norm[k] = 0;
    if (lane == 0) {
      c->status = 2;
      c->reason = WHY_DEFECT;
      c->events += 1;
    }
    return;
  }
  const bool first = c->iter == 0;
  unsigned long long imax = 0, nmax = 0, tmax = 0;
  int open = 0;
  for (int k = lane; k < g.ncase; k += kLanes) {
    const double norm = __longlong_as_double(static_cast<long long>(g.norm[k]));
    if (first) g.initial[k] = norm;
    const double initial = g.initial[k];
    const double tol = fmax(g.atol, g.rtol * initial);
    const unsigned long long bi = static_cast<unsigned long long>(__double_as_longlong(initial));
    const unsigned long long bn = static_cast<unsigned long long>(__double_as_longlong(norm));
    const unsigned long long bt = static_cast<unsigned long long>(__double_as_longlong(tol));
    imax = bi > imax ? bi : imax;
    nmax = bn > nmax ? bn : nmax;
    tmax = bt > tmax ? bt : tmax;
    if (!isfinite(tol) || norm > tol) open = 1;
    g.norm[k] = 0;
  }
  imax = wave_max(imax);
  nmax = wave_max(nmax);
  tmax = wave_max(tmax);
  const bool converged = !__any(open);
  if (converged) {
    const double tend = c->tend;
    for (int k = lane; k < g.ncase; k += kLanes) {
      const double d = g.dt[k];
      if (d <= 0.0) continue;
      const double t = g.t[k];
      g.t[k] = d == tend - t ? tend : t + d;
    }
  }
  if (lane == 0) {
    c->initial_defect = __longlong_as_double(static_cast<long long>(imax));
    c->final_defect = __longlong_as_double(static_cast<long long>(nmax));
    c->tolerance = __longlong_as_double(static_cast<long long>(tmax));
    c->checks += 1;
    if (converged) {
      c->cur = c->qbuf;
      c->nsteps += 1;
      c->steps += 1;
      c->phase = PH_BOUND;
    } else if (c->iter >= g.maxiter) {
      c->status = 4;
      c->reason = WHY_MAXITER;
    } else {
      c->phase = PH_TRIAL;
      c->pretries = 0;
    }
    c->events += 1;
  }
}

// Pseudo-time candidate: trial = Q + omega (dt R - (Q - U)), omega = cfl* / (cfl* + dt rate(Q))
// for the cases that move, then the admissibility of every case's candidate cells.
__global__ __launch_bounds__(kThreads) void k_be_trial(Grid g) {
  const Control* c = g.ctrl;
  if (c->status != RUNNING || c->phase != PH_TRIAL) return;
  const int cur = c->cur, qb = c->qbuf, tb = 3 - cur - qb;
  const double pseudo = c->pseudo;
  const double* u = g.B[cur];
  const double* Q = g.B[qb];
  double* trial = g.B[tb];
  const size_t inner = static_cast<size_t>(g.nx) * g.ny;
  const size_t pl = g.plane;
  bool bad = false;
  for (size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; q < inner * g.ncase;
       q += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t cs = q / inner, ij = q - cs * inner;
    const int j = static_cast<int>(ij / g.nx), i = static_cast<int>(ij - static_cast<size_t>(j) * g.nx);
    const size_t kc = static_cast<size_t>(j + kGhosts) * g.mx + i + kGhosts;
    if (g.has_cut && g.vol[kc] <= 0.0) continue;
    const size_t k = cs * 4 * pl + kc;
    const double d = g.dt[cs];
    if (d > 0.0) {
      double rate;
      if (!g.has_cut) {
        rate = wave_bound(Q[k], Q[k + pl], Q[k + 2 * pl], Q[k + 3 * pl], g, g.dx[i + kGhosts], g.dy[j + kGhosts]);
      } else {
        // implicit_spectral_rate of a cell of a cut grid (src/implicit2d.h:33-43)
        const double r = Q[k], mu = Q[k + pl], mv = Q[k + 2 * pl], E = Q[k + 3 * pl];
        const double u = mu / r, v = mv / r;
        const double p = pressure(r, mu, mv, E, g.gamma);
        const double a = sqrt(g.gamma * p / r);
        const size_t fe = static_cast<size_t>(j) * (g.nx + 1) + i, fn = static_cast<size_t>(j) * g.nx + i;
        const double wx = g.wax[kc], wy = g.way[kc];
        const double wall = sqrt(wx * wx + wy * wy);
        rate = ((fabs(u) + a) * (g.axl[fe + 1] + g.axl[fe]) + (fabs(v) + a) * (g.ayl[fn + g.nx] + g.ayl[fn]) +
                fabs(u * wx + v * wy) + a * wall) /
               (2.0 * g.vol[kc]);
      }
      const double omega = pseudo / (pseudo + d * rate);
      for (int v = 0; v < 4; ++v) {
        const size_t kv = k + v * pl;
        trial[kv] = Q[kv] + omega * (d * g.R[kv] - (Q[kv] - u[kv]));
      }
    }
    if (inadmissible(trial[k], trial[k + pl], trial[k + 2 * pl], trial[k + 3 * pl], g.gamma)) bad = true;
  }
  if (__any(bad) && threadIdx.x % kLanes == 0) atomicOr(&g.ctrl->bad, 1);
}

// Candidate accepted as the next iterate, or the pseudo CFL halved (12 retries, then 2).
__global__ __launch_bounds__(kLanes) void k_be_check(Grid g) {
  Control* c = g.ctrl;
  if (threadIdx.x != 0 || c->status != RUNNING || c->phase != PH_TRIAL) return;
  if (c->bad) {
    c->bad = 0;
    c->iretries += 1;
    c->pretries += 1;
    if (c->pretries > kMaxRetries) {
      c->status = 2;
      c->reason = WHY_PSEUDO;
    } else {
      c->pseudo *= 0.5;
    }
  } else {
    c->qbuf = 3 - c->cur - c->qbuf;
    c->updates += 1;
    c->iter += 1;
    c->phase = PH_RES;
  }
  c->events += 1;
}

// ------------------------------------------------------------------------------- SRD --
// State redistribution of the stage's candidate (src/euler2d_cuda.cu:316-368, 461-468), when it is
// admissible: neighbourhood averages over the members, then every rewritten cell sums its own
// value (regular cells) and the averages of the neighbourhoods it belongs to, in ascending
// order, and divides by its count.  MODE picks the candidate buffer as the stage kernels do.
// A trial is rejected when the first stage's candidate is inadmissible in any cell (tested
// before the redistribution, src/euler2d_cuda.cu:648-651); on a cut grid that happens in the
// smallest cut cells, several times per step, until the step has been halved enough.  This
// kernel forms the first-stage candidate of every cut cell (and case) alone, with k_stage<1>'s
// formulas in its order (so the same bits: neighbours are interior, the body keeps three cells
// from every side), and raises the trial's flag if one is inadmissible; the stage kernels then
// return at once.  The decision is the one the whole-grid pass would take.
__global__ __launch_bounds__(kThreads) void k_stage1_precheck(Grid g) {
  const Control* c = g.ctrl;
  if (__hip_atomic_load(&c->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != RUNNING ||
      __hip_atomic_load(&c->bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || c->pre_done)
    return;
  const double* in = g.B[c->cur];
  const size_t plane = g.plane;
  const size_t n = static_cast<size_t>(g.ncase) * g.n_pre;
  bool bad = false;
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t icase = t / g.n_pre;
    const int kc = g.pre_cell[t - icase * g.n_pre];
    const int jr = kc / g.mx - kGhosts, i = kc % g.mx - kGhosts;
    const double* u = in + icase * 4 * plane;
    const double mco = g.mco[icase], dt = g.dt[icase];
    const auto prim_at = [&](int k) {
      return g.vol[k] <= 0.0 ? Prim{1.0, 0.0, 0.0, 1.0}
                             : primitive(u[k], u[k + plane], u[k + 2 * plane], u[k + 3 * plane], g.gamma);
    };
    const Prim w = prim_at(kc);
    const double ax = g.axl[static_cast<size_t>(jr) * (g.nx + 1) + i];
    const double ae = g.axl[static_cast<size_t>(jr) * (g.nx + 1) + i + 1];
    const double as = g.ayl[static_cast<size_t>(jr) * g.nx + i];
    const double an = g.ayl[static_cast<size_t>(jr + 1) * g.nx + i];
    Flux fw{0.0, 0.0, 0.0, 0.0}, fe{0.0, 0.0, 0.0, 0.0}, gs{0.0, 0.0, 0.0, 0.0}, gn{0.0, 0.0, 0.0, 0.0};
    if (ax > 0.0) fw = face_flux<true>(g, mco, prim_at(kc - 1), w);
    if (ae > 0.0) fe = face_flux<true>(g, mco, w, prim_at(kc + 1));
    if (as > 0.0) gs = face_flux<false>(g, mco, prim_at(kc - g.mx), w);
    if (an > 0.0) gn = face_flux<false>(g, mco, w, prim_at(kc + g.mx));
    const Flux xres{ae * fe.m - ax * fw.m, ae * fe.x - ax * fw.x, ae * fe.y - ax * fw.y, ae * fe.e - ax * fw.e};
    const double iv = 1.0 / g.vol[kc], pw = w.p;
    const double r0 = -iv * (xres.m + an * gn.m - as * gs.m);
    double r1 = -iv * (xres.x + an * gn.x - as * gs.x);
    double r2 = -iv * (xres.y + an * gn.y - as * gs.y);
    const double r3 = -iv * (xres.e + an * gn.e - as * gs.e);
    r1 -= iv * pw * g.wax[kc];
    r2 -= iv * pw * g.way[kc];
    constexpr double A = 0.0, B = 1.0;  // the first stage's combination, as k_stage<1> forms it
    const double q0 = u[kc], q1 = u[kc + plane], q2 = u[kc + 2 * plane], q3 = u[kc + 3 * plane];
    const double n0 = A * q0 + B * (q0 + dt * r0);
    const double n1 = A * q1 + B * (q1 + dt * r1);
    const double n2 = A * q2 + B * (q2 + dt * r2);
    const double n3 = A * q3 + B * (q3 + dt * r3);
    if (inadmissible(n0, n1, n2, n3, g.gamma)) bad = true;
  }
  if (__any(bad) && threadIdx.x % kLanes == 0) atomicOr(&g.ctrl->bad, 1);
}

// After a pre-check (one wave): a failed candidate is rejected at once, as the trial's commit
// would reject it (halve every positive step, count the retry); a passing one ends the
// pre-checks of this trial.  A trial runs kPrechecks of these pairs before its first stage, so
// a step that needs several halvings costs them here instead of a whole-grid pass each.
__global__ __launch_bounds__(kLanes) void k_precheck_reject(Grid g) {
  Control* c = g.ctrl;
  if (c->status != RUNNING || c->pre_done) return;
  if (__hip_atomic_load(&c->bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    reject(g);
  else if (threadIdx.x == 0)
    c->pre_done = 1;
}

template <int MODE>
__device__ __forceinline__ double* stage_out(const Grid& g, int cur) {
  return MODE == 2 ? g.B[2] : g.B[cur ^ 1];
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_srd_avg(Grid g) {
  const Control* c = g.ctrl;
  if (__hip_atomic_load(&c->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != RUNNING ||
      __hip_atomic_load(&c->bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    return;
  const double* u = stage_out<MODE>(g, c->cur);
  const size_t n = static_cast<size_t>(g.ncase) * g.nsmall;
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t icase = t / g.nsmall;
    const int q = static_cast<int>(t - icase * g.nsmall);
    const double* uc = u + icase * 4 * g.plane;
    for (int v = 0; v < 4; ++v) {
      double sum = 0.0;
      for (int m = g.mem_off[q]; m < g.mem_off[q + 1]; ++m) {
        const int r = g.mem_idx[m];
        sum += g.vol[r] / g.nbrcount[r] * uc[v * g.plane + r];
      }
      g.Qh[t * 4 + v] = sum / g.svhat[q];
    }
  }
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_srd_apply(Grid g) {
  Control* c = g.ctrl;
  if (__hip_atomic_load(&c->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != RUNNING ||
      __hip_atomic_load(&c->bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    return;
  double* u = stage_out<MODE>(g, c->cur);
  const size_t n = static_cast<size_t>(g.ncase) * g.n_srd;
  const int lane = threadIdx.x % kLanes;
  bool bad = false;
  // Wave-uniform iterations (64 consecutive items per wave), so that the wave can reduce the
  // stage-3 bounds of one case to a single atomic.
  for (size_t base = blockIdx.x * static_cast<size_t>(blockDim.x) + (threadIdx.x - lane); base < n;
       base += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t t = base + lane;
    const bool valid = t < n;
    const size_t icase = valid ? t / g.n_srd : 0;
    unsigned long long bits = 0;
    if (valid) {
      const int e = static_cast<int>(t - icase * g.n_srd);
      const int k = g.srd_cell[e];
      double* uc = u + icase * 4 * g.plane;
      const double* qh = g.Qh + icase * g.nsmall * 4;
      double x[4];
      for (int v = 0; v < 4; ++v) {
        double acc = g.smallpos[k] < 0.0 ? uc[v * g.plane + k] : 0.0;
        for (int m = g.grp_off[e]; m < g.grp_off[e + 1]; ++m) acc += qh[g.grp_idx[m] * 4 + v];
        x[v] = acc / g.nbrcount[k];
        uc[v * g.plane + k] = x[v];
      }
      if (inadmissible(x[0], x[1], x[2], x[3], g.gamma)) bad = true;
      if (MODE == 3) {
        const int j = k / g.mx - kGhosts, i = k % g.mx - kGhosts;
        const double b = wave_bound(x[0], x[1], x[2], x[3], g, g.dx[i + kGhosts], g.dy[j + kGhosts]);
        bits = static_cast<unsigned long long>(__double_as_longlong(b));
      }
    }
    if (MODE == 3) {
      const size_t first = __shfl(icase, 0, kLanes);
      if (__all(!valid || icase == first)) {  // one case in the wave: one atomic
        const unsigned long long m = wave_max(bits);
        if (lane == 0 && m) atomicMax(&g.smax[first], m);
      } else if (bits) {
        atomicMax(&g.smax[icase], bits);
      }
    }
  }
  if (__any(bad) && lane == 0) atomicOr(&c->bad, 1);
}

__global__ void k_zero(double* p, size_t n) {
  for (size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; q < n;
       q += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[q] = 0.0;
}

}  // namespace ugrid

// =============================================================================================
// Host side
// =============================================================================================
using namespace ugrid;

struct ugrid_session {
  int device = 0;
  Grid g{};
  ugrid_params par{};
  size_t states = 0;      // doubles per state array (ghosts included)
  int launch_blocks = 0;  // padded to a multiple of 8
  double* R = nullptr;       // residual output of ugrid_rhs (allocated on first use)
  double* rhs_in = nullptr;  // input staging of ugrid_rhs (allocated on first use)
  std::vector<double> mco;   // host copy of the reference Mach numbers
  Control* host_ctrl = nullptr;  // pinned mirror of the control record
  hipStream_t stream = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_trials = 0;
  std::vector<double> uploaded;  // the state as uploaded (its ghost cells are the download's
                                 // ghosts until a step begins)
  std::vector<void*> owned;
};

namespace {

int fail(char* err, size_t len, const std::string& message) {
  if (err && len) {
    const size_t k = message.size() < len - 1 ? message.size() : len - 1;
    std::memcpy(err, message.data(), k);
    err[k] = 0;
  }
  return 3;
}

bool ok(hipError_t e, char* err, size_t len, const char* what) {
  if (e == hipSuccess) return true;
  fail(err, len, std::string(what) + ": " + hipGetErrorString(e));
  return false;
}

template <class T>
bool device_alloc(ugrid_session* s, T** p, size_t count, char* err, size_t len) {
  void* q = nullptr;
  if (!ok(hipMalloc(&q, (count ? count : 1) * sizeof(T)), err, len, "hipMalloc")) return false;
  s->owned.push_back(q);
  *p = static_cast<T*>(q);
  return true;
}

void release(ugrid_session* s) {
  if (!s) return;
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  for (void* p : s->owned) (void)hipFree(p);
  if (s->host_ctrl) (void)hipHostFree(s->host_ctrl);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// Fill the two ghost layers of every plane from the interior as the reference's boundary kernel
// does (src/euler2d_cuda.cu:153-180): both indices mapped, normal momentum negated on reflection.
void fill_ghosts(const Grid& g, double* u) {
  const int ng = kGhosts;
  for (int c = 0; c < g.ncase; ++c)
    for (int v = 0; v < 4; ++v) {
      double* q = u + (static_cast<size_t>(c) * 4 + v) * g.plane;
      for (int j = -ng; j < g.ny + ng; ++j)
        for (int i = -ng; i < g.nx + ng; ++i) {
          if (i >= 0 && i < g.nx && j >= 0 && j < g.ny) continue;
          int ii = i, jj = j;
          double sign = 1.0;
          if (ii < 0) {
            if (g.bc[0] == BC_PERIODIC) ii += g.nx;
            else if (g.bc[0] == BC_REFLECT) { ii = -ii - 1; if (v == 1) sign = -sign; }
            else ii = 0;
          } else if (ii >= g.nx) {
            if (g.bc[1] == BC_PERIODIC) ii -= g.nx;
            else if (g.bc[1] == BC_REFLECT) { ii = 2 * g.nx - 1 - ii; if (v == 1) sign = -sign; }
            else ii = g.nx - 1;
          }
          if (jj < 0) {
            if (g.bc[2] == BC_PERIODIC) jj += g.ny;
            else if (g.bc[2] == BC_REFLECT) { jj = -jj - 1; if (v == 2) sign = -sign; }
            else jj = 0;
          } else if (jj >= g.ny) {
            if (g.bc[3] == BC_PERIODIC) jj -= g.ny;
            else if (g.bc[3] == BC_REFLECT) { jj = 2 * g.ny - 1 - jj; if (v == 2) sign = -sign; }
            else jj = g.ny - 1;
          }
          q[static_cast<size_t>(j + ng) * g.mx + i + ng] = sign * q[static_cast<size_t>(jj + ng) * g.mx + ii + ng];
        }
    }
}

// Blocks per case of k_bound: about 2048 blocks over the batch, at most one per 256 cells.
void launch_bound(ugrid_session* s) {
  const Grid& g = s->g;
  const long long inner = static_cast<long long>(g.nx) * g.ny;
  const long long cap = (inner + kThreads - 1) / kThreads;
  const int per_case = static_cast<int>(std::max<long long>(1, std::min<long long>(cap, 2048 / g.ncase)));
  hipLaunchKernelGGL(k_bound, dim3(per_case * g.ncase), dim3(kThreads), 0, s->stream, g, per_case, -1);
}

int grid_blocks(const Grid& g, size_t items) {
  return static_cast<int>(std::max<size_t>(1, std::min<size_t>((items + kThreads - 1) / kThreads, 8192)));
}

// A stage kernel in the regular or the cut-grid instantiation.
template <int MODE>
void launch_stage(ugrid_session* s, hipStream_t stream, const double* in = nullptr, double* out = nullptr) {
  if (s->g.has_cut)
    hipLaunchKernelGGL((k_stage<MODE, true>), dim3(s->launch_blocks), dim3(kThreads), 0, stream, s->g, in, out);
  else
    hipLaunchKernelGGL((k_stage<MODE, false>), dim3(s->launch_blocks), dim3(kThreads), 0, stream, s->g, in, out);
}

// A stage kernel followed, on a cut grid with small cells, by the redistribution of its candidate.
template <int MODE>
void launch_stage_srd(ugrid_session* s, hipStream_t stream) {
  const Grid& g = s->g;
  if (MODE == 1 && g.has_cut && g.n_pre > 0) {
    // kPrechecks pre-checks, each failed one rejected at once; the last one's failure is left to
    // the trial's commit (the stage kernels then return at once).
    const dim3 pre_blocks(grid_blocks(g, static_cast<size_t>(g.ncase) * g.n_pre));
    for (int r = 0; r < kPrechecks; ++r) {
      hipLaunchKernelGGL(k_stage1_precheck, pre_blocks, dim3(kThreads), 0, stream, g);
      if (r + 1 < kPrechecks) hipLaunchKernelGGL(k_precheck_reject, dim3(1), dim3(kLanes), 0, stream, g);
    }
  }
  launch_stage<MODE>(s, stream);
  if (!g.has_cut || g.nsmall == 0) return;
  hipLaunchKernelGGL(k_srd_avg<MODE>, dim3(grid_blocks(g, static_cast<size_t>(g.ncase) * g.nsmall)), dim3(kThreads), 0,
                     stream, g);
  hipLaunchKernelGGL(k_srd_apply<MODE>, dim3(grid_blocks(g, static_cast<size_t>(g.ncase) * g.n_srd)),
                     dim3(kThreads), 0, stream, g);
}

// One backward-Euler replay: every phase kernel once, in phase order.
int launch_be_round(ugrid_session* s, hipStream_t stream) {
  const Grid& g = s->g;
  const long long inner = static_cast<long long>(g.nx) * g.ny;
  const int per_case = static_cast<int>(
      std::max<long long>(1, std::min<long long>((inner + kThreads - 1) / kThreads, 2048 / g.ncase)));
  const size_t cells = static_cast<size_t>(inner) * g.ncase;
  hipLaunchKernelGGL(k_bound, dim3(per_case * g.ncase), dim3(kThreads), 0, stream, g, per_case, PH_BOUND);
  hipLaunchKernelGGL(k_be_open, dim3(1), dim3(kLanes), 0, stream, g);
  hipLaunchKernelGGL(k_be_copy, dim3(grid_blocks(g, 4 * cells)), dim3(kThreads), 0, stream, g);
  launch_stage<4>(s, stream);
  hipLaunchKernelGGL(k_be_decide, dim3(1), dim3(kLanes), 0, stream, g);
  hipLaunchKernelGGL(k_be_trial, dim3(grid_blocks(g, cells)), dim3(kThreads), 0, stream, g);
  hipLaunchKernelGGL(k_be_check, dim3(1), dim3(kLanes), 0, stream, g);
  return 7;
}

int launch_trial(ugrid_session* s, hipStream_t stream) {
  launch_stage_srd<1>(s, stream);
  launch_stage_srd<2>(s, stream);
  launch_stage_srd<3>(s, stream);
  hipLaunchKernelGGL(k_commit, dim3(1), dim3(kLanes), 0, stream, s->g);
  return 4;
}

bool build_graph(ugrid_session* s, int trials, char* err, size_t len) {
  if (s->exec) {
    (void)hipGraphExecDestroy(s->exec);
    s->exec = nullptr;
  }
  hipGraph_t graph = nullptr;
  if (!ok(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), err, len, "begin capture")) return false;
  for (int k = 0; k < trials; ++k) {
    if (s->par.time_integrator == 1)
      launch_be_round(s, s->stream);
    else
      launch_trial(s, s->stream);
  }
  const hipError_t launched = hipGetLastError();
  const hipError_t ended = hipStreamEndCapture(s->stream, &graph);
  if (!ok(launched, err, len, "trial launch") || !ok(ended, err, len, "end capture")) {
    if (graph) (void)hipGraphDestroy(graph);
    return false;
  }
  const hipError_t inst = hipGraphInstantiate(&s->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (!ok(inst, err, len, "graph instantiate")) return false;
  s->graph_trials = trials;
  return true;
}

// Cut-cell arrays of the device: per-cell data, apertures, and the SRD lists (members of each
// neighbourhood; the cells SRD rewrites, each with its neighbourhoods in ascending order).
bool upload_cut(ugrid_session* s, const ugrid_cut_view& c, const double* dx, const double* dy, char* err, size_t len) {
  Grid& g = s->g;
  const size_t cells = g.plane, nfx = static_cast<size_t>(g.nx + 1) * g.ny, nfy = static_cast<size_t>(g.nx) * (g.ny + 1);
  std::vector<int> mem_off(1, 0), mem_idx;
  std::vector<std::vector<int>> groups(cells);
  for (int q = 0; q < c.nsmall; ++q) {
    const int count = static_cast<int>(c.scnt[q]);
    if (count < 1 || count > 25) {
      fail(err, len, "ugrid_create: invalid SRD neighbourhood size");
      return false;
    }
    for (int m = 0; m < count; ++m) {
      const double r = c.smem[25 * static_cast<size_t>(q) + m];
      if (!(r >= 0.0) || r >= static_cast<double>(cells)) {
        fail(err, len, "ugrid_create: invalid SRD member");
        return false;
      }
      mem_idx.push_back(static_cast<int>(r));
      groups[static_cast<size_t>(r)].push_back(q);
    }
    mem_off.push_back(static_cast<int>(mem_idx.size()));
  }
  std::vector<int> srd_cell, grp_off(1, 0), grp_idx;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const size_t k = static_cast<size_t>(j + kGhosts) * g.mx + i + kGhosts;
      if (c.vol[k] <= 0.0 || (c.smallpos[k] < 0.0 && c.nbrcount[k] == 1.0)) continue;
      srd_cell.push_back(static_cast<int>(k));
      grp_idx.insert(grp_idx.end(), groups[k].begin(), groups[k].end());
      grp_off.push_back(static_cast<int>(grp_idx.size()));
    }
  // Cell codes of the stage kernels (see Grid::code).  An aperture is coded 1 only when it equals
  // the full width bit for bit, 0 only when it is +0, so the kernels read the same values.
  std::vector<uint8_t> code(cells, 0);
  const auto face_state = [](double a, double full) -> unsigned {
    if (a == 0.0 && !std::signbit(a)) return 0u;
    return a == full ? 1u : 2u;
  };
  for (int jj = 0; jj < g.my; ++jj)
    for (int ii = 0; ii < g.mx; ++ii) {
      const size_t k = static_cast<size_t>(jj) * g.mx + ii;
      const int j = jj - kGhosts, i = ii - kGhosts;
      unsigned v = 0;
      if (c.vol[k] <= 0.0) v |= 1u;
      if (c.flag[k] != 0.0) v |= 2u;
      if (c.smallpos[k] >= 0.0 || c.nbrcount[k] != 1.0) v |= 4u;
      if (j >= 0 && j < g.ny && i >= 0 && i <= g.nx)
        v |= face_state(c.axl[static_cast<size_t>(j) * (g.nx + 1) + i], dy[jj]) << 3;
      if (j >= 0 && j <= g.ny && i >= 0 && i < g.nx)
        v |= face_state(c.ayl[static_cast<size_t>(j) * g.nx + i], dx[ii]) << 5;
      code[k] = static_cast<uint8_t>(v);
    }
  std::vector<int> pre_cell;  // interior cut cells, for the first-stage precheck
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const size_t k = static_cast<size_t>(j + kGhosts) * g.mx + i + kGhosts;
      if (c.flag[k] != 0.0 && c.vol[k] > 0.0) pre_cell.push_back(static_cast<int>(k));
    }
  g.has_cut = 1;
  g.nsmall = c.nsmall;
  g.n_srd = static_cast<int>(srd_cell.size());
  g.n_pre = static_cast<int>(pre_cell.size());
  double *vol, *flag, *axl, *ayl, *wax, *way, *nbr, *pos, *svhat, *qh;
  int *moff, *midx, *scell, *goff, *gidx, *pcell;
  uint8_t* dcode;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream), err, len, "upload cut geometry");
  };
  const size_t nq = static_cast<size_t>(std::max(c.nsmall, 1));
  if (!device_alloc(s, &vol, cells, err, len) || !device_alloc(s, &flag, cells, err, len) ||
      !device_alloc(s, &axl, nfx, err, len) || !device_alloc(s, &ayl, nfy, err, len) ||
      !device_alloc(s, &wax, cells, err, len) || !device_alloc(s, &way, cells, err, len) ||
      !device_alloc(s, &nbr, cells, err, len) || !device_alloc(s, &pos, cells, err, len) ||
      !device_alloc(s, &svhat, nq, err, len) || !device_alloc(s, &qh, nq * 4 * g.ncase, err, len) ||
      !device_alloc(s, &moff, mem_off.size(), err, len) || !device_alloc(s, &midx, std::max<size_t>(mem_idx.size(), 1), err, len) ||
      !device_alloc(s, &scell, std::max<size_t>(srd_cell.size(), 1), err, len) ||
      !device_alloc(s, &goff, grp_off.size(), err, len) || !device_alloc(s, &gidx, std::max<size_t>(grp_idx.size(), 1), err, len) ||
      !device_alloc(s, &pcell, std::max<size_t>(pre_cell.size(), 1), err, len) ||
      !device_alloc(s, &dcode, cells, err, len))
    return false;
  const size_t d = sizeof(double), n = sizeof(int);
  if (!up(vol, c.vol, cells * d) || !up(flag, c.flag, cells * d) || !up(axl, c.axl, nfx * d) ||
      !up(ayl, c.ayl, nfy * d) || !up(wax, c.wax, cells * d) || !up(way, c.way, cells * d) ||
      !up(nbr, c.nbrcount, cells * d) || !up(pos, c.smallpos, cells * d) ||
      (c.nsmall > 0 && !up(svhat, c.svhat, c.nsmall * d)) || !up(moff, mem_off.data(), mem_off.size() * n) ||
      (!mem_idx.empty() && !up(midx, mem_idx.data(), mem_idx.size() * n)) ||
      (!srd_cell.empty() && !up(scell, srd_cell.data(), srd_cell.size() * n)) ||
      !up(goff, grp_off.data(), grp_off.size() * n) || (!grp_idx.empty() && !up(gidx, grp_idx.data(), grp_idx.size() * n)) ||
      (!pre_cell.empty() && !up(pcell, pre_cell.data(), pre_cell.size() * n)) || !up(dcode, code.data(), cells) ||
      !ok(hipStreamSynchronize(s->stream), err, len, "upload cut geometry"))
    return false;
  g.vol = vol;
  g.flag = flag;
  g.axl = axl;
  g.ayl = ayl;
  g.wax = wax;
  g.way = way;
  g.nbrcount = nbr;
  g.smallpos = pos;
  g.svhat = svhat;
  g.Qh = qh;
  g.mem_off = moff;
  g.mem_idx = midx;
  g.srd_cell = scell;
  g.grp_off = goff;
  g.grp_idx = gidx;
  g.pre_cell = pcell;
  g.code = dcode;
  return true;
}

// Rows per strip: 32, halved (down to 4) while the launch would hold fewer than 16 waves per
// CU (256 CUs), and never more than the grid has.
int choose_rows(int nx, int ny, int ncase) {
  const long long sx = (nx + kStripCells - 1) / kStripCells;
  int rows = 32;
  while (rows > 4 && sx * ((ny + rows - 1) / rows) * ncase < 16LL * 256) rows /= 2;
  return rows < ny ? rows : ny;
}

}  // namespace

extern "C" int ugrid_create(const ugrid_setup* setup, const double* u, uint64_t u_len, const double* t,
                            const double* dt, ugrid_session** out, char* err, size_t err_len) {
  if (!out) return fail(err, err_len, "ugrid_create: null output handle");
  *out = nullptr;
  if (!setup || !u || !setup->dx || !setup->dy) return fail(err, err_len, "ugrid_create: null argument");
  const ugrid_params& p = setup->params;
  if (setup->nx < 1 || setup->ny < 1 || setup->ncase < 1)
    return fail(err, err_len, "ugrid_create: nx, ny and ncase must be positive");
  if (p.time_integrator != 0 && p.time_integrator != 1)
    return fail(err, err_len, "ugrid_create: time_integrator must be 0 (SSP-RK3) or 1 (backward Euler)");
  if (p.limiter != 3)
    return fail(err, err_len,
                "ugrid_create: the device backend requires the first-order limiter (3); refusing to change the method "
                "implicitly");
  for (int k = 0; k < 4; ++k)
    if (p.bc[k] < 0 || p.bc[k] > 2) return fail(err, err_len, "ugrid_create: boundary kinds are 0, 1 or 2");
  if ((p.bc[0] == BC_PERIODIC) != (p.bc[1] == BC_PERIODIC) || (p.bc[2] == BC_PERIODIC) != (p.bc[3] == BC_PERIODIC))
    return fail(err, err_len, "ugrid_create: periodic boundary must be set on both sides of an axis");
  if (setup->cut) {
    const ugrid_cut_view& c = *setup->cut;
    if (c.nbad_nbhd != 0)
      return fail(err, err_len, "ugrid_create: the device backend refuses " + std::to_string(c.nbad_nbhd) +
                                    " invalid SRD neighbourhoods");
    if (c.nsmall < 0 || !c.vol || !c.flag || !c.axl || !c.ayl || !c.wax || !c.way || !c.nbrcount || !c.smallpos ||
        (c.nsmall > 0 && (!c.scnt || !c.smem || !c.svhat)))
      return fail(err, err_len, "ugrid_create: incomplete cut-cell geometry");
    // The stage kernels take a ghost cell's state from the interior cell its boundary maps to,
    // which holds for the ghosts' own fluid volumes only while the body keeps three cells
    // clear of every side (as the geometry builders enforce).
    const int nx = setup->nx, ny = setup->ny;
    for (int j = -kGhosts; j < ny + kGhosts; ++j)
      for (int i = -kGhosts; i < nx + kGhosts; ++i) {
        const size_t k = static_cast<size_t>(j + kGhosts) * (nx + 2 * kGhosts) + (i + kGhosts);
        const bool near = i < 3 || i >= nx - 3 || j < 3 || j >= ny - 3;
        if (near && (c.flag[k] != 0.0 || !(c.vol[k] > 0.0)))
          return fail(err, err_len, "ugrid_create: cut-cell geometry within three cells of the domain boundary (cell " +
                                        std::to_string(i) + "," + std::to_string(j) + ")");
      }
    const double cells = static_cast<double>(nx + 2 * kGhosts) * (ny + 2 * kGhosts);
    for (int q = 0; q < c.nsmall; ++q) {
      const double count = c.scnt[q];
      if (!(count >= 1.0 && count <= 25.0) || count != static_cast<int>(count))
        return fail(err, err_len, "ugrid_create: invalid SRD neighbourhood size");
      for (int m = 0; m < static_cast<int>(count); ++m) {
        const double r = c.smem[25 * static_cast<size_t>(q) + m];
        if (!(r >= 0.0 && r < cells) || r != static_cast<double>(static_cast<long long>(r)))
          return fail(err, err_len, "ugrid_create: invalid SRD member");
      }
    }
  }
  const long long mx = setup->nx + 2LL * kGhosts, my = setup->ny + 2LL * kGhosts;
  if (mx * my > (1LL << 40) / setup->ncase) return fail(err, err_len, "ugrid_create: grid too large");
  const size_t states = static_cast<size_t>(setup->ncase) * 4 * static_cast<size_t>(mx * my);
  if (u_len != states) return fail(err, err_len, "ugrid_create: u length must be ncase * 4 * (nx + 4) * (ny + 4)");

  ugrid_session* s = new ugrid_session();
  s->par = p;
  s->states = states;
  Grid& g = s->g;
  g.nx = setup->nx;
  g.ny = setup->ny;
  g.mx = static_cast<int>(mx);
  g.my = static_cast<int>(my);
  g.ncase = setup->ncase;
  g.plane = static_cast<size_t>(mx * my);
  for (int k = 0; k < 4; ++k) g.bc[k] = p.bc[k];
  g.ausm = p.flux == 1;
  g.gamma = p.gamma;
  g.cfl = p.cfl;
  g.dt_fixed = p.dt_fixed;
  g.pseudo0 = p.pseudo_cfl;
  g.rtol = p.implicit_rtol;
  g.atol = p.implicit_atol;
  g.maxiter = p.implicit_maxiter;
  g.rows = choose_rows(g.nx, g.ny, g.ncase);
  g.strips_x = (g.nx + kStripCells - 1) / kStripCells;
  g.strips_y = (g.ny + g.rows - 1) / g.rows;
  const long long strips = static_cast<long long>(g.strips_x) * g.strips_y * g.ncase;
  if (strips > (1LL << 30)) {
    release(s);
    return fail(err, err_len, "ugrid_create: grid too large");
  }
  g.strips = static_cast<int>(strips);
  g.blocks = (g.strips + kWaves - 1) / kWaves;
  s->launch_blocks = (g.blocks + kXcds - 1) / kXcds * kXcds;

  std::vector<double> mco(g.ncase, p.ausm_mco);
  if (setup->mco)
    for (int c = 0; c < g.ncase; ++c) mco[c] = setup->mco[c];
  std::vector<double> t0(g.ncase, 0.0), dt0(g.ncase, 0.0);
  if (t)
    for (int c = 0; c < g.ncase; ++c) t0[c] = t[c];
  if (dt)
    for (int c = 0; c < g.ncase; ++c) dt0[c] = dt[c];

  double *dxd = nullptr, *dyd = nullptr, *mcod = nullptr;
  if (!ok(hipGetDevice(&s->device), err, err_len, "hipGetDevice") ||
      !ok(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), err, err_len, "hipStreamCreate") ||
      !ok(hipHostMalloc(reinterpret_cast<void**>(&s->host_ctrl), sizeof(Control), hipHostMallocDefault), err, err_len,
          "hipHostMalloc") ||
      !device_alloc(s, &g.B[0], states, err, err_len) || !device_alloc(s, &g.B[1], states, err, err_len) ||
      !device_alloc(s, &g.B[2], states, err, err_len) || !device_alloc(s, &dxd, g.mx, err, err_len) ||
      (p.time_integrator == 1 && !device_alloc(s, &g.R, states, err, err_len)) ||
      !device_alloc(s, &g.norm, g.ncase, err, err_len) || !device_alloc(s, &g.initial, g.ncase, err, err_len) ||
      !device_alloc(s, &dyd, g.my, err, err_len) || !device_alloc(s, &mcod, g.ncase, err, err_len) ||
      !device_alloc(s, &g.t, g.ncase, err, err_len) || !device_alloc(s, &g.dt, g.ncase, err, err_len) ||
      !device_alloc(s, &g.dt_out, g.ncase, err, err_len) || !device_alloc(s, &g.smax, g.ncase, err, err_len) ||
      !device_alloc(s, &g.ctrl, 1, err, err_len)) {
    release(s);
    return 3;
  }
  g.dx = dxd;
  g.dy = dyd;
  g.mco = mcod;
  Control c0{};
  c0.status = 0;
  const hipMemcpyKind h2d = hipMemcpyHostToDevice;
  const size_t bytes = states * sizeof(double);
  // Every buffer that can come to hold the accepted state starts as the upload, so its ghost
  // cells are the entry values (the device never writes ghost cells).
  if (!ok(hipMemcpyAsync(g.B[0], u, bytes, h2d, s->stream), err, err_len, "upload state") ||
      !ok(hipMemcpyAsync(g.B[1], u, bytes, h2d, s->stream), err, err_len, "upload state") ||
      !ok(hipMemcpyAsync(g.B[2], u, bytes, h2d, s->stream), err, err_len, "upload state") ||
      !ok(hipMemsetAsync(g.norm, 0, g.ncase * sizeof(unsigned long long), s->stream), err, err_len, "clear norms") ||
      !ok(hipMemcpyAsync(dxd, setup->dx, g.mx * sizeof(double), h2d, s->stream), err, err_len, "upload dx") ||
      !ok(hipMemcpyAsync(dyd, setup->dy, g.my * sizeof(double), h2d, s->stream), err, err_len, "upload dy") ||
      !ok(hipMemcpyAsync(mcod, mco.data(), g.ncase * sizeof(double), h2d, s->stream), err, err_len, "upload mco") ||
      !ok(hipMemcpyAsync(g.t, t0.data(), g.ncase * sizeof(double), h2d, s->stream), err, err_len, "upload t") ||
      !ok(hipMemcpyAsync(g.dt, dt0.data(), g.ncase * sizeof(double), h2d, s->stream), err, err_len, "upload dt") ||
      !ok(hipMemcpyAsync(g.dt_out, dt0.data(), g.ncase * sizeof(double), h2d, s->stream), err, err_len,
          "upload dt") ||
      !ok(hipMemsetAsync(g.smax, 0, g.ncase * sizeof(unsigned long long), s->stream), err, err_len, "clear bound") ||
      !ok(hipMemcpyAsync(g.ctrl, &c0, sizeof(Control), h2d, s->stream), err, err_len, "upload control") ||
      !ok(hipStreamSynchronize(s->stream), err, err_len, "upload")) {
    release(s);
    return 3;
  }
  s->uploaded.assign(u, u + states);
  s->mco = mco;
  if (setup->cut && !upload_cut(s, *setup->cut, setup->dx, setup->dy, err, err_len)) {
    release(s);
    return 3;
  }
  const int trials = g.nx * static_cast<long long>(g.ny) * g.ncase >= (1 << 22) ? 4 : 16;
  if (!build_graph(s, trials, err, err_len)) {
    release(s);
    return 3;
  }
  *out = s;
  return 0;
}

extern "C" int ugrid_run(ugrid_session* s, double tend, int64_t maxsteps, ugrid_meta* meta, char* err,
                         size_t err_len) {
  if (!s) return fail(err, err_len, "ugrid_run: null session");
  const Grid& g = s->g;
  const bool implicit = s->par.time_integrator == 1;
  if (implicit) {
    // implicit2d_validate (src/implicit2d.cpp:11-35): controls, times and reference Mach numbers.
    const ugrid_params& p = s->par;
    std::vector<double> t(g.ncase);
    if (!ok(hipMemcpyAsync(t.data(), g.t, g.ncase * sizeof(double), hipMemcpyDeviceToHost, s->stream), err, err_len,
            "download t") ||
        !ok(hipStreamSynchronize(s->stream), err, err_len, "download t"))
      return 3;
    bool valid = std::isfinite(tend) && maxsteps >= 0 && p.gamma > 1.0 && std::isfinite(p.gamma) && p.cfl > 0.0 &&
                 std::isfinite(p.cfl) && p.dt_fixed >= 0.0 && std::isfinite(p.dt_fixed) && p.pseudo_cfl > 0.0 &&
                 std::isfinite(p.pseudo_cfl) && p.implicit_maxiter >= 1 && p.implicit_rtol >= 0.0 &&
                 std::isfinite(p.implicit_rtol) && p.implicit_atol > 0.0 && std::isfinite(p.implicit_atol);
    for (int c = 0; c < g.ncase; ++c)
      valid = valid && std::isfinite(t[c]) && s->mco[c] > 0.0 && std::isfinite(s->mco[c]);
    if (!valid) return fail(err, err_len, "invalid backward-Euler controls");
    const size_t cells = static_cast<size_t>(g.nx) * g.ny * g.ncase;
    hipLaunchKernelGGL(k_check_state, dim3(grid_blocks(g, cells)), dim3(kThreads), 0, s->stream, g);
    hipLaunchKernelGGL(k_be_start, dim3(1), dim3(kLanes), 0, s->stream, g, tend, static_cast<long long>(maxsteps));
  } else {
    launch_bound(s);
    hipLaunchKernelGGL(k_open, dim3(1), dim3(kThreads), 0, s->stream, g, tend, static_cast<long long>(maxsteps));
  }
  if (!ok(hipGetLastError(), err, err_len, "run start")) return 3;
  // Every replay of a running batch takes at least one decision (a commit, a rejection, a
  // phase change), so it must move the event count; a replay that does not is reported, not
  // waited on.
  long long progress = -1;
  for (;;) {
    if (!ok(hipMemcpyAsync(s->host_ctrl, g.ctrl, sizeof(Control), hipMemcpyDeviceToHost, s->stream), err, err_len,
            "control copy") ||
        !ok(hipStreamSynchronize(s->stream), err, err_len, "trial"))
      return 3;
    if (s->host_ctrl->status != RUNNING) break;
    const long long now = s->host_ctrl->events;
    if (now == progress) return fail(err, err_len, "ugrid_run: device step control made no progress");
    progress = now;
    if (!ok(hipGraphLaunch(s->exec, s->stream), err, err_len, "trial graph")) return 3;
  }
  const Control& c = *s->host_ctrl;
  if (meta) {
    meta->nsteps = c.nsteps;
    meta->nreject = c.nreject;
  }
  char line[256];
  line[0] = 0;
  if (implicit) {
    if (c.reason == WHY_INITIAL)
      std::snprintf(line, sizeof(line), "backward Euler: invalid initial fluid state");
    else if (c.reason == WHY_PSEUDO)
      std::snprintf(line, sizeof(line), "backward Euler: nonphysical pseudo candidate; physical step not accepted");
    else if (c.reason == WHY_DEFECT)
      std::snprintf(line, sizeof(line), "backward Euler: non-finite defect; physical step not accepted");
    else if (c.reason == WHY_MAXITER)
      std::snprintf(line, sizeof(line),
                    "backward Euler not converged: step=%lld updates=%d initial=%.6e defect=%.6e tolerance=%.6e; "
                    "physical step not accepted",
                    c.nsteps, c.iter, c.initial_defect, c.final_defect, c.tolerance);
    else if (c.reason == WHY_DT)
      std::snprintf(line, sizeof(line), "backward Euler: physical dt is invalid or cannot advance time");
  } else if (c.status == 2) {
    std::snprintf(line, sizeof(line), "non-physical RK candidate after %d retries at accepted step %lld", c.retries,
                  c.nsteps);
  }
  if (line[0]) fail(err, err_len, line);
  return c.status;
}

extern "C" int ugrid_download(ugrid_session* s, double* u, uint64_t u_len, double* t, double* dt, char* err,
                              size_t err_len) {
  if (!s) return fail(err, err_len, "ugrid_download: null session");
  if (u && u_len != s->states) return fail(err, err_len, "ugrid_download: u length mismatch");
  const Grid& g = s->g;
  Control c{};
  if (!ok(hipMemcpyAsync(&c, g.ctrl, sizeof(Control), hipMemcpyDeviceToHost, s->stream), err, err_len, "control") ||
      !ok(hipStreamSynchronize(s->stream), err, err_len, "control"))
    return 3;
  const size_t bytes = s->states * sizeof(double);
  if (u) {
    if (!ok(hipMemcpyAsync(u, g.B[c.cur], bytes, hipMemcpyDeviceToHost, s->stream), err, err_len, "download") ||
        !ok(hipStreamSynchronize(s->stream), err, err_len, "download"))
      return 3;
    std::vector<double> other;
    const double* ghosts = s->uploaded.data();
    if (c.ghost_from == 1) {
      fill_ghosts(g, u);
      ghosts = nullptr;
    } else if (c.ghost_from == 2) {
      other.resize(s->states);
      if (!ok(hipMemcpyAsync(other.data(), g.B[c.cur ^ 1], bytes, hipMemcpyDeviceToHost, s->stream), err, err_len,
              "download") ||
          !ok(hipStreamSynchronize(s->stream), err, err_len, "download"))
        return 3;
      fill_ghosts(g, other.data());
      ghosts = other.data();
    }
    if (ghosts)
      for (size_t cv = 0; cv < static_cast<size_t>(g.ncase) * 4; ++cv)
        for (int j = 0; j < g.my; ++j)
          for (int i = 0; i < g.mx; ++i) {
            if (i >= kGhosts && i < g.nx + kGhosts && j >= kGhosts && j < g.ny + kGhosts) continue;
            const size_t k = cv * g.plane + static_cast<size_t>(j) * g.mx + i;
            u[k] = ghosts[k];
          }
  }
  if (t && !ok(hipMemcpyAsync(t, g.t, g.ncase * sizeof(double), hipMemcpyDeviceToHost, s->stream), err, err_len,
                "download t"))
    return 3;
  if (dt && !ok(hipMemcpyAsync(dt, g.dt_out, g.ncase * sizeof(double), hipMemcpyDeviceToHost, s->stream), err,
                 err_len, "download dt"))
    return 3;
  return ok(hipStreamSynchronize(s->stream), err, err_len, "download") ? 0 : 3;
}

extern "C" int ugrid_rhs(ugrid_session* s, const double* u, uint64_t u_len, double* r, char* err, size_t err_len) {
  if (!s || !u || !r) return fail(err, err_len, "ugrid_rhs: null argument");
  if (u_len != s->states) return fail(err, err_len, "ugrid_rhs: u length mismatch");
  const size_t bytes = s->states * sizeof(double);
  if (!s->R && !device_alloc(s, &s->R, s->states, err, err_len)) return 3;
  if (!s->rhs_in && !device_alloc(s, &s->rhs_in, s->states, err, err_len)) return 3;
  hipLaunchKernelGGL(k_zero, dim3(1024), dim3(kThreads), 0, s->stream, s->R, s->states);
  if (!ok(hipMemcpyAsync(s->rhs_in, u, bytes, hipMemcpyHostToDevice, s->stream), err, err_len, "upload state"))
    return 3;
  launch_stage<0>(s, s->stream, s->rhs_in, s->R);
  if (!ok(hipGetLastError(), err, err_len, "residual launch") ||
      !ok(hipMemcpyAsync(r, s->R, bytes, hipMemcpyDeviceToHost, s->stream), err, err_len, "download residual") ||
      !ok(hipStreamSynchronize(s->stream), err, err_len, "residual"))
    return 3;
  return 0;
}

extern "C" int ugrid_implicit_stats(const ugrid_session* s, ugrid_implicit_info* out) {
  if (!s || !out) return 1;
  Control c{};
  if (hipMemcpyAsync(&c, s->g.ctrl, sizeof(Control), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess)
    return 1;
  out->updates = c.updates;
  out->retries = c.iretries;
  out->checks = c.checks;
  out->reason = c.reason;
  out->initial_defect = c.initial_defect;
  out->final_defect = c.final_defect;
  out->tolerance = c.tolerance;
  return 0;
}

extern "C" int ugrid_profile(ugrid_session* s, int32_t trials, double* stage_us, char* err, size_t err_len) {
  if (!s || !stage_us || trials < 1) return fail(err, err_len, "ugrid_profile: invalid argument");
  if (s->par.time_integrator != 0) return fail(err, err_len, "ugrid_profile: SSP-RK3 sessions only");
  const Grid& g = s->g;
  launch_bound(s);
  hipLaunchKernelGGL(k_open, dim3(1), dim3(kThreads), 0, s->stream, g, HUGE_VAL, (1LL << 62));
  std::vector<hipEvent_t> ev(4 * static_cast<size_t>(trials), nullptr);
  bool good = true;
  for (auto& e : ev) good = good && hipEventCreate(&e) == hipSuccess;
  for (int k = 0; good && k < trials; ++k) {
    hipEvent_t* e = &ev[4 * static_cast<size_t>(k)];
    (void)hipEventRecord(e[0], s->stream);
    launch_stage_srd<1>(s, s->stream);
    (void)hipEventRecord(e[1], s->stream);
    launch_stage_srd<2>(s, s->stream);
    (void)hipEventRecord(e[2], s->stream);
    launch_stage_srd<3>(s, s->stream);
    (void)hipEventRecord(e[3], s->stream);
    hipLaunchKernelGGL(k_commit, dim3(1), dim3(kLanes), 0, s->stream, g);
  }
  good = good && ok(hipGetLastError(), err, err_len, "profile launch") &&
         ok(hipStreamSynchronize(s->stream), err, err_len, "profile");
  double sum[3] = {0, 0, 0};
  for (int k = 0; good && k < trials; ++k)
    for (int q = 0; q < 3; ++q) {
      float ms = 0;
      good = good && ok(hipEventElapsedTime(&ms, ev[4 * static_cast<size_t>(k) + q], ev[4 * static_cast<size_t>(k) + q + 1]),
                        err, err_len, "profile events");
      sum[q] += ms;
    }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (!good) return 3;
  for (int q = 0; q < 3; ++q) stage_us[q] = 1e3 * sum[q] / trials;
  return 0;
}

extern "C" int ugrid_device_available(char* reason, size_t reason_len) {
  int count = 0;
  const hipError_t e = hipGetDeviceCount(&count);
  char line[256];
  if (e != hipSuccess || count <= 0) {
    std::snprintf(line, sizeof(line), "hipGetDeviceCount: %s (devices=%d)", hipGetErrorString(e), count);
    fail(reason, reason_len, line);
    return 0;
  }
  int device = 0;
  hipDeviceProp_t p{};
  if (hipGetDevice(&device) != hipSuccess || hipGetDeviceProperties(&p, device) != hipSuccess) {
    fail(reason, reason_len, "hipGetDeviceProperties failed");
    return 0;
  }
  std::snprintf(line, sizeof(line), "%s, %s", p.name, p.gcnArchName);
  fail(reason, reason_len, line);
  return 1;
}

extern "C" int ugrid_get_launch(const ugrid_session* s, ugrid_launch_info* out) {
  if (!s || !out) return 1;
  out->rows = s->g.rows;
  out->strips = s->g.strips;
  out->blocks = s->launch_blocks;
  out->threads = kThreads;
  return 0;
}

extern "C" void ugrid_destroy(ugrid_session* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  (void)hipStreamSynchronize(s->stream);
  release(s);
}
