// amr_kernels.hip -- the explicit cut-cell SSP-RK3 trial as gfx950 kernels.
//
// One trial (reference launch_trial, src/static_amr2d_cuda.cu:241-268) is, per RHS pass:
//   k_grad     cells: viscosity, eddy viscosity, LSQ velocity gradient, SA-neg gradients and
//              point model (reference :71-91, src/static_amr2d_sa_cuda.cuh:2-11)
//   k_face     faces then walls: boundary state, HLLC / AUSM+-up, viscous face flux, SA face
//              flux; no-slip cut-wall shear (reference :92-143, src/static_amr2d_sa_cuda.cuh:12-29)
//   k_collect  cells: CSR gather r = -(sum F + p n l - wall shear)/V; the base pass also forms
//              the local CFL / viscous / SA step and reduces it (wave shuffles + LDS + one 64-bit
//              atomicMin per block); stage passes apply the SSP-RK3 combination and convert the
//              candidate to primitives in the same kernel (reference :144-163,188-194)
// plus k_base_prim, k_stage0 (dt choice + first stage, reference :170-191), the SRD pair
// k_group / k_redist (reference :195-207), the SRD-grouped viscous limit (reference :164-176)
// and k_commit (accept / reject / begin, reference :54-59,212-216).
// Every kernel serves every mesh of the arena; a block reads its mesh from a block table and
// skips the whole tile when that mesh is not running (block-uniform branch).
#include <cfloat>

#include "amr_arena.h"

namespace hipamr {
namespace {

constexpr int kMaxTile = 256;
constexpr unsigned long long kInfBits = 0x7ff0000000000000ull;

enum : int { PASS_BASE = 0, PASS_STAGE = 1 };

__device__ __forceinline__ bool running(const Control* c) { return c->active && !c->fatal; }
__device__ __forceinline__ bool healthy(const Control* c) { return running(c) && !c->bad; }
__device__ __forceinline__ bool pass_on(const Control* c, int pass) {
  return healthy(c) && (pass == PASS_STAGE || !c->r0_valid);
}
// Admissibility bits: 1 = the base pass failed (fatal 2 at the dt choice), 2 = a candidate
// stage failed (the trial is rejected).  Separate bits keep the dt choice race-free.
enum : int { BAD_BASE = 1, BAD_STAGE = 2 };
__device__ __forceinline__ void flag_bad(Control* c, int bit) { atomicOr(&c->bad, bit); }

// U -> W = (rho, u, v, p) and nutilde (computed in FP64 from the FP64 state, stored as T);
// false if not admissible (reference primitive rules, src/static_amr2d.cpp:143-154).
template <class T>
__device__ __forceinline__ bool to_primitive(const double* u, int nc, int k, T gamma, bool sa, T* w, T* phi) {
  const double rho = u[k], mx = u[nc + k], my = u[2 * nc + k], energy = u[3 * nc + k];
  const double ir = 1.0 / rho;
  const double vx = mx * ir, vy = my * ir;
  const double p = (static_cast<double>(gamma) - 1.0) * (energy - 0.5 * (mx * vx + my * vy));
  bool ok = isfinite(rho) && rho > 0 && isfinite(mx) && isfinite(my) && isfinite(energy) && isfinite(p) && p > 0;
  w[0] = static_cast<T>(rho);
  w[1] = static_cast<T>(vx);
  w[2] = static_cast<T>(vy);
  w[3] = static_cast<T>(p);
  if (sa) {
    const double ph = u[4 * nc + k] * ir;
    *phi = static_cast<T>(ph);
    ok = ok && isfinite(ph);
  }
  return ok && w[0] > T(0) && w[3] > T(0) && isfinite(w[0]) && isfinite(w[3]);
}

template <class T>
__device__ __forceinline__ void store_primitive(const ArenaPtrs<T>& a, int k, const T* w, T phi, bool sa) {
  a.W[k] = w[0];
  a.W[a.nc + k] = w[1];
  a.W[2 * a.nc + k] = w[2];
  a.W[3 * a.nc + k] = w[3];
  if (sa) a.PHI[k] = phi;
}

// Minimum over the block of a double (wave64 xor-shuffles, then one LDS slot per wave).
// All threads of the block must call it; the result is valid in thread 0.
__device__ __forceinline__ double block_min(double v) {
  __shared__ double part[kMaxTile / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) part[wave] = v;
  __syncthreads();
  double r = part[0];
  if (threadIdx.x == 0)
    for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) r = fmin(r, part[w]);
  return r;
}

__device__ __forceinline__ void publish_dt(Control* c, double local) {
  const double m = block_min(local);
  if (threadIdx.x == 0 && m < INFINITY) atomicMin(&c->dt_min, static_cast<unsigned long long>(__double_as_longlong(m)));
}

// ------------------------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kMaxTile) void k_base_prim(ArenaPtrs<T> a) {
  const int m = a.cell_block_mesh[blockIdx.x];
  Control* c = a.ctrl + m;
  if (!pass_on(c, PASS_BASE) || c->w_valid) return;
  const MeshDesc<T>& d = a.desc[m];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.cell_begin + d.n_cells) return;
  T w[4], phi = T(0);
  const bool ok = to_primitive(a.U[c->cur], a.nc, k, d.gamma, d.sa != 0, w, &phi);
  if (!ok) {
    flag_bad(c, BAD_BASE);
    atomicExch(&c->fatal, 2);  // the accepted state itself is not admissible
    return;
  }
  store_primitive(a, k, w, phi, d.sa != 0);
}

// ------------------------------------------------------------------------------------------
template <class T, int PASS>
__global__ __launch_bounds__(kMaxTile) void k_grad(ArenaPtrs<T> a) {
  const int m = a.cell_block_mesh[blockIdx.x];
  Control* c = a.ctrl + m;
  if (!pass_on(c, PASS)) return;
  const MeshDesc<T>& d = a.desc[m];
  if (!d.laminar) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.cell_begin + d.n_cells) return;
  const int nc = a.nc;
  const bool sa = d.sa != 0;
  const T rho = a.W[k], u = a.W[nc + k], v = a.W[2 * nc + k], p = a.W[3 * nc + k];
  const T mu = viscosity(d.transport, p / (rho * d.transport.gas_constant));
  const int bit = PASS == PASS_BASE ? BAD_BASE : BAD_STAGE;
  if (!(mu > T(0)) || !isfinite(mu)) {
    flag_bad(c, bit);
    return;
  }
  a.MU[k] = mu;
  const T phi = sa ? a.PHI[k] : T(0);
  if (sa) a.EDDY[k] = rho * sa_eddy(phi, mu / rho);
  T g0 = 0, g1 = 0, g2 = 0, g3 = 0, sx = 0, sy = 0, rx = 0, ry = 0;
  const int end = a.term_off[k + 1];
  for (int j = a.term_off[k]; j < end; ++j) {
    const int src = a.term_src[j];
    const T wx = a.term_wx[j], wy = a.term_wy[j];
    T uu = T(0), vv = T(0), ph = T(0), rr = rho;
    if (src >= 0) {
      uu = a.W[nc + src];
      vv = a.W[2 * nc + src];
      if (sa) {
        ph = a.PHI[src];
        rr = a.W[src];
      }
    } else if (src <= -2) {
      const Side<T>& s = d.side[-2 - src];
      uu = s.wall_velocity[0];
      vv = s.wall_velocity[1];
    }
    const T du = uu - u, dv = vv - v;
    g0 += wx * du;
    g1 += wy * du;
    g2 += wx * dv;
    g3 += wy * dv;
    if (sa) {
      sx += wx * (ph - phi);
      sy += wy * (ph - phi);
      rx += wx * (rr - rho);
      ry += wy * (rr - rho);
    }
  }
  a.GRAD[k] = g0;
  a.GRAD[nc + k] = g1;
  a.GRAD[2 * nc + k] = g2;
  a.GRAD[3 * nc + k] = g3;
  if (sa) {
    a.SGRAD[k] = sx;
    a.SGRAD[nc + k] = sy;
    const SaPoint<T> pt = sa_point(phi, mu / rho, a.wall_dist[k], fabs(g2 - g1));
    const T source = rho * (pt.source + T(SaConst<T>::cb2 / SaConst<T>::sigma) * (sx * sx + sy * sy)) -
                     pt.diffusion * (rx * sx + ry * sy);
    a.SDIFF[k] = pt.diffusion;
    a.SSRC[k] = source;
    a.SRATE[k] = pt.source_rate;
    if (!(pt.diffusion > T(0)) || !isfinite(source) || !isfinite(pt.source_rate)) flag_bad(c, bit);
  }
}

// ------------------------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ Prim<T> load_prim(const ArenaPtrs<T>& a, int k) {
  return Prim<T>{a.W[k], a.W[a.nc + k], a.W[2 * a.nc + k], a.W[3 * a.nc + k]};
}

// One face: boundary state, inviscid flux in the face frame, viscous and SA face fluxes.
// Everything stays in registers: the face-normal component is chosen with selects on the
// axis, never by indexing a local array (which would put it in scratch memory).
template <class T, int PASS>
__device__ __forceinline__ void face_point(const ArenaPtrs<T>& a, const MeshDesc<T>& d, Control* c, int q) {
  const int nc = a.nc, nf = a.nf, bit = PASS == PASS_BASE ? BAD_BASE : BAD_STAGE;
  const int l = a.face_l[q], r = a.face_r[q], info = a.face_info[q];
  const int axis = info & 1, side = (info >> 1) - 1;
  const bool xaxis = axis == 0;
  const T len = a.face_len[q];
  const int inside = l >= 0 ? l : r;
  Prim<T> wl = load_prim(a, l >= 0 ? l : inside);
  Prim<T> wr = load_prim(a, r >= 0 ? r : inside);
  int kind = -1;
  if (side >= 0) {
    kind = d.side[side].kind;
    const bool ok = l < 0 ? side_state(d.side[side], d.gamma, side, wr, wl) : side_state(d.side[side], d.gamma, side, wl, wr);
    if (!ok) {
      flag_bad(c, bit);
      return;
    }
  }
  // Face frame: n = normal velocity, t = tangential velocity.
  const T nL = xaxis ? wl.u : wl.v, tL = xaxis ? wl.v : wl.u;
  const T nR = xaxis ? wr.u : wr.v, tR = xaxis ? wr.v : wr.u;
  T f[4];
  if (kind == KIND_NOSLIP) {
    f[0] = T(0);
    f[1] = a.W[3 * nc + inside];
    f[2] = T(0);
    f[3] = T(0);
  } else if (d.flux == 0) {
    hllc_flux(wl.r, nL, tL, wl.p, wr.r, nR, tR, wr.p, d.gamma, f);
  } else {
    ausm_flux(wl.r, nL, tL, wl.p, wr.r, nR, tR, wr.p, d.gamma, d.mco, f);
  }
  if (!isfinite(f[0]) || !isfinite(f[1]) || !isfinite(f[2]) || !isfinite(f[3])) {
    flag_bad(c, bit);
    return;
  }
  T F0 = f[0] * len, Fx = (xaxis ? f[1] : f[2]) * len, Fy = (xaxis ? f[2] : f[1]) * len, F3 = f[3] * len;
  if (d.laminar) {
    // Viscous face flux (reference viscous_face, src/static_amr2d_cuda.cu:92-107).
    const bool sa = d.sa != 0;
    const int k = inside;
    T g0 = a.GRAD[k], g1 = a.GRAD[nc + k], g2 = a.GRAD[2 * nc + k], g3 = a.GRAD[3 * nc + k];
    const T rho_k = a.W[k], u_k = a.W[nc + k], v_k = a.W[2 * nc + k];
    const T eddy_k = sa ? a.EDDY[k] : T(0);
    T mu = a.MU[k] + eddy_k, uf = u_k, vf = v_k;
    T nu = mu / rho_k;
    const T dist = a.face_dist[q];
    if (side < 0) {
      const T dx = a.face_dx[q], dy = a.face_dy[q], ex = a.face_ex[q], ey = a.face_ey[q];
      const T gl[4] = {a.GRAD[l], a.GRAD[nc + l], a.GRAD[2 * nc + l], a.GRAD[3 * nc + l]};
      const T gr[4] = {a.GRAD[r], a.GRAD[nc + r], a.GRAD[2 * nc + r], a.GRAD[3 * nc + r]};
      T g[4];
      corrected_gradient(gl, gr, dx, dy, wr.u - wl.u, wr.v - wl.v, g);
      g0 = g[0];
      g1 = g[1];
      g2 = g[2];
      g3 = g[3];
      uf = T(0.5) * (wl.u + wr.u) + g0 * ex + g1 * ey;
      vf = T(0.5) * (wl.v + wr.v) + g2 * ex + g3 * ey;
      const T temp = T(0.5) * (wl.p / wl.r + wr.p / wr.r) / d.transport.gas_constant;
      const T el = sa ? a.EDDY[l] : T(0), er = sa ? a.EDDY[r] : T(0);
      mu = viscosity(d.transport, temp) + T(0.5) * (el + er);
      nu = fmax((a.MU[l] + el) / wl.r, (a.MU[r] + er) / wr.r);
    } else {
      const T sign = (side & 1) ? T(1) : T(-1);
      if (kind == KIND_NOSLIP) {
        mu = a.MU[k];
        uf = d.side[side].wall_velocity[0];
        vf = d.side[side].wall_velocity[1];
        const T du = (uf - u_k) / dist, dv = (vf - v_k) / dist;
        const T nx = xaxis ? sign : T(0), ny = xaxis ? T(0) : sign;
        g0 = nx * du;
        g1 = ny * du;
        g2 = nx * dv;
        g3 = ny * dv;
      } else if (kind == KIND_REFLECT) {
        g1 = T(0);
        g2 = T(0);
        if (xaxis) uf = T(0);
        else vf = T(0);
      } else {
        if (xaxis) {
          g0 = T(0);
          g2 = T(0);
        } else {
          g1 = T(0);
          g3 = T(0);
        }
        const T ex = a.face_ex[q], ey = a.face_ey[q];
        uf += g0 * ex + g1 * ey;
        vf += g2 * ex + g3 * ey;
      }
    }
    const T g[4] = {g0, g1, g2, g3};
    T fv[4];
    stress_flux(mu, g, xaxis ? T(1) : T(0), xaxis ? T(0) : T(1), uf, vf, fv);
    if (!isfinite(fv[1]) || !isfinite(fv[2]) || !isfinite(fv[3])) flag_bad(c, bit);
    Fx -= fv[1] * len;
    Fy -= fv[2] * len;
    F3 -= fv[3] * len;
    if (PASS == PASS_BASE) a.FCOEF[q] = T(4.0 / 3.0) * nu / dist;
    if (sa) {
      // SA face flux: upwind nutilde times the mass flux, minus corrected-gradient diffusion
      // (reference sa_face_wall_point, src/static_amr2d_sa_cuda.cuh:16-25).
      const T mass = F0;
      const int ir = r >= 0 ? r : inside;  // always a valid cell index
      const T pl = l >= 0 ? a.PHI[inside] : d.transport.sa_nu_inf;
      const T pr = r >= 0 ? a.PHI[ir] : d.transport.sa_nu_inf;
      T diff = T(0), fc = T(0);
      if (side < 0) {
        const T dx = a.face_dx[q], dy = a.face_dy[q];
        const T mx = T(0.5) * (a.SGRAD[l] + a.SGRAD[r]), my = T(0.5) * (a.SGRAD[nc + l] + a.SGRAD[nc + r]);
        const T corr = ((pr - pl) - mx * dx - my * dy) / (dx * dx + dy * dy);
        const T gn = xaxis ? mx + corr * dx : my + corr * dy;
        const T coeff = T(0.5) * (wl.r * a.SDIFF[l] + wr.r * a.SDIFF[r]);
        diff = coeff * gn * len;
        fc = coeff / dist * len;
      } else {
        const bool inflow = l < 0 ? mass > T(0) : mass < T(0);
        const T sign = (side & 1) ? T(1) : T(-1);
        if (kind == KIND_NOSLIP || inflow) {
          const T wallv = kind == KIND_NOSLIP ? T(0) : d.transport.sa_nu_inf;
          const T coeff = kind == KIND_NOSLIP ? a.MU[inside] * T(1.0 / SaConst<T>::sigma) : rho_k * a.SDIFF[inside];
          diff = sign * coeff * (wallv - a.PHI[inside]) / dist * len;
          fc = coeff / dist * len;
        }
      }
      const T sflux = mass * (mass >= T(0) ? pl : pr) - diff;
      a.SFLUX[q] = sflux;
      if (PASS == PASS_BASE) a.SFCOEF[q] = fc;
      if (!isfinite(sflux)) flag_bad(c, bit);
    }
  }
  a.FLUX[q] = F0;
  a.FLUX[nf + q] = Fx;
  a.FLUX[2 * nf + q] = Fy;
  a.FLUX[3 * nf + q] = F3;
  if (PASS == PASS_BASE)
    a.SPEED[q] = fmax(fabs(nL) + sqrt(d.gamma * wl.p / wl.r), fabs(nR) + sqrt(d.gamma * wr.p / wr.r));
}

// No-slip shear on a cut-wall fragment from its donor cell (reference viscous_wall_point,
// src/static_amr2d_cuda.cu:108-113, and the SA wall part of sa_face_wall_point).
template <class T, int PASS>
__device__ __forceinline__ void wall_point(const ArenaPtrs<T>& a, const MeshDesc<T>& d, Control* c, int j) {
  const int nc = a.nc, k = a.wall_donor[j];
  const T nx = a.wall_nx[j], ny = a.wall_ny[j], len = a.wall_len[j], dd = a.wall_ddist[j];
  const T rho = a.W[k], mu = a.MU[k];
  const T du = -a.W[nc + k] / dd, dv = -a.W[2 * nc + k] / dd;
  const T g[4] = {nx * du, ny * du, nx * dv, ny * dv};
  T fv[4];
  stress_flux(mu, g, nx, ny, T(0), T(0), fv);
  if (!isfinite(fv[1]) || !isfinite(fv[2])) flag_bad(c, PASS == PASS_BASE ? BAD_BASE : BAD_STAGE);
  a.WFLUX[j] = fv[1] * len;
  a.WFLUX[a.nw + j] = fv[2] * len;
  if (PASS == PASS_BASE) {
    const int o = a.wall_owner[j];
    a.WCOEF[j] = T(4.0 / 3.0) * fmax(mu / rho, a.MU[o] / a.W[o]) / dd;
  }
  if (d.sa) {
    const T coeff = mu * T(1.0 / SaConst<T>::sigma) / dd * len;
    if (PASS == PASS_BASE) a.SWCOEF[j] = coeff;
    a.SWFLUX[j] = -coeff * a.PHI[k];
  }
}

template <class T, int PASS>
__global__ __launch_bounds__(kMaxTile) void k_face(ArenaPtrs<T> a) {
  const int b = blockIdx.x;
  const int m = a.face_block_mesh[b];
  Control* c = a.ctrl + m;
  if (!pass_on(c, PASS)) return;
  const MeshDesc<T>& d = a.desc[m];
  if (b < a.n_face_blocks) {
    const int q = b * blockDim.x + threadIdx.x;
    if (q < d.face_begin + d.n_faces) face_point<T, PASS>(a, d, c, q);
  } else if (d.laminar) {
    const int j = (b - a.n_face_blocks) * blockDim.x + threadIdx.x;
    if (j < d.wall_begin + d.n_walls) wall_point<T, PASS>(a, d, c, j);
  }
}

// ------------------------------------------------------------------------------------------
// STAGE 0: base residual R0 and local dt.  1, 2: second and third SSP-RK3 combinations.
template <class T, int STAGE>
__global__ __launch_bounds__(kMaxTile) void k_collect(ArenaPtrs<T> a) {
  const int m = a.cell_block_mesh[blockIdx.x];
  Control* c = a.ctrl + m;
  if (!pass_on(c, STAGE == 0 ? PASS_BASE : PASS_STAGE)) return;
  const MeshDesc<T>& d = a.desc[m];
  const int nc = a.nc, nf = a.nf, nw = a.nw;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = k < d.cell_begin + d.n_cells;
  const bool sa = d.sa != 0, laminar = d.laminar != 0;
  double local_dt = INFINITY;
  if (valid) {
    T s0 = 0, s1 = 0, s2 = 0, s3 = 0, spectral = 0, diffusion = 0, sa_net = 0, sa_spec = 0;
    const int end = a.csr_off[k + 1];
    for (int j = a.csr_off[k]; j < end; ++j) {
      const int sq = a.csr_face[j];
      const int q = sq >= 0 ? sq : ~sq;
      const T sgn = sq >= 0 ? T(1) : T(-1);
      s0 += sgn * a.FLUX[q];
      s1 += sgn * a.FLUX[nf + q];
      s2 += sgn * a.FLUX[2 * nf + q];
      s3 += sgn * a.FLUX[3 * nf + q];
      if (sa) sa_net += sgn * a.SFLUX[q];
      if (STAGE == 0) {
        const T len = a.face_len[q];
        spectral += a.SPEED[q] * len;
        if (laminar) diffusion += a.FCOEF[q] * len;
        if (sa) sa_spec += a.SFCOEF[q];
      }
    }
    const T rho = a.W[k], p = a.W[3 * nc + k];
    const int w0 = a.wall_first[k], w1 = w0 + a.wall_cnt[k];
    if (w1 > w0) {
      const T snd = sqrt(d.gamma * p / rho);
      for (int j = w0; j < w1; ++j) {
        const T len = a.wall_len[j];
        s1 += p * a.wall_nx[j] * len;
        s2 += p * a.wall_ny[j] * len;
        if (STAGE == 0) spectral += snd * len;
        if (laminar) {
          s1 -= a.WFLUX[j];
          s2 -= a.WFLUX[nw + j];
          if (STAGE == 0) diffusion += a.WCOEF[j] * len;
        }
        if (sa) {
          sa_net -= a.SWFLUX[j];
          if (STAGE == 0) sa_spec += a.SWCOEF[j];
        }
      }
    }
    const T iv = a.inv_vol[k];
    const T r0 = -s0 * iv, r1 = -s1 * iv, r2 = -s2 * iv, r3 = -s3 * iv;
    T r4 = T(0);
    if (sa) {
      r4 = -sa_net * iv + a.SSRC[k];
      if (!isfinite(r4)) flag_bad(c, STAGE == 0 ? BAD_BASE : BAD_STAGE);
    }
    if (STAGE == 0) {
      a.R0[k] = r0;
      a.R0[nc + k] = r1;
      a.R0[2 * nc + k] = r2;
      a.R0[3 * nc + k] = r3;
      if (sa) a.R0[4 * nc + k] = r4;
      T dt = spectral > T(0) ? d.cfl * a.dt_area[k] / spectral : T(INFINITY);
      if (laminar) {
        if (d.visc_mode == VISC_VOLUME) {
          if (diffusion > T(0)) dt = fmin(dt, d.transport.viscous_cfl * a.vol[k] / diffusion);
        } else if (d.visc_mode == VISC_SRD_RATE) {
          const T rate = diffusion * iv;
          if (rate > T(0)) dt = fmin(dt, d.transport.viscous_cfl / rate);
        } else {
          a.CELL_DT[k] = dt;
          a.DRATE[k] = diffusion * iv;
        }
      }
      if (sa) {
        const T rate = T(2) * sa_spec / (rho * a.vol[k]) + a.SRATE[k];
        if (rate > T(0)) dt = fmin(dt, d.transport.sa_cfl / rate);
      }
      local_dt = static_cast<double>(dt);
    } else {
      // SSP-RK3: u2 = 3/4 u + 1/4 (u1 + dt r)  or  u1 = 1/3 u + 2/3 (u2 + dt r).
      const double dt = c->trial_dt;
      const int cur = c->cur;
      const double* u = a.U[cur];
      double* out = STAGE == 1 ? a.U2 : a.U[cur ^ 1];
      const double* prev = STAGE == 1 ? a.U[cur ^ 1] : a.U2;
      const double wu = STAGE == 1 ? 0.75 : 1.0 / 3.0, wp = STAGE == 1 ? 0.25 : 2.0 / 3.0;
      out[k] = wu * u[k] + wp * (prev[k] + dt * static_cast<double>(r0));
      out[nc + k] = wu * u[nc + k] + wp * (prev[nc + k] + dt * static_cast<double>(r1));
      out[2 * nc + k] = wu * u[2 * nc + k] + wp * (prev[2 * nc + k] + dt * static_cast<double>(r2));
      out[3 * nc + k] = wu * u[3 * nc + k] + wp * (prev[3 * nc + k] + dt * static_cast<double>(r3));
      if (sa) out[4 * nc + k] = wu * u[4 * nc + k] + wp * (prev[4 * nc + k] + dt * static_cast<double>(r4));
      if (!d.small) {
        T w[4], phi = T(0);
        if (!to_primitive(out, nc, k, d.gamma, sa, w, &phi)) flag_bad(c, BAD_STAGE);
        else store_primitive(a, k, w, phi, sa);
      }
    }
  }
  if (STAGE == 0 && d.visc_mode != VISC_SRD_GROUP) publish_dt(c, local_dt);
}

// SRD-grouped viscous limit: group means of the diffusion rate, then the per-cell limit
// (reference diffusion_group_kernel and choose_dt SRD branch, src/static_amr2d_cuda.cu:164-176).
template <class T>
__global__ __launch_bounds__(kMaxTile) void k_diffusion_group(ArenaPtrs<T> a) {
  const int m = a.cell_block_mesh[blockIdx.x];
  Control* c = a.ctrl + m;
  if (!pass_on(c, PASS_BASE)) return;
  const MeshDesc<T>& d = a.desc[m];
  if (d.visc_mode != VISC_SRD_GROUP) return;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= d.cell_begin + d.n_cells) return;
  T sum = 0;
  for (int j = a.grp_off[g]; j < a.grp_off[g + 1]; ++j) sum += a.grp_w[j] * a.DRATE[a.grp_mem[j]];
  a.DGROUP[g] = sum;
}

template <class T>
__global__ __launch_bounds__(kMaxTile) void k_dt_group(ArenaPtrs<T> a) {
  const int m = a.cell_block_mesh[blockIdx.x];
  Control* c = a.ctrl + m;
  if (!pass_on(c, PASS_BASE)) return;
  const MeshDesc<T>& d = a.desc[m];
  if (d.visc_mode != VISC_SRD_GROUP) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double local = INFINITY;
  if (k < d.cell_begin + d.n_cells) {
    T sum = 0;
    for (int j = a.mem_off[k]; j < a.mem_off[k + 1]; ++j) sum += a.DGROUP[a.mem_grp[j]];
    const T rate = sum * a.mem_inv[k];
    T dt = a.CELL_DT[k];
    if (rate > T(0)) dt = fmin(dt, d.transport.viscous_cfl / rate);
    local = static_cast<double>(dt);
  }
  publish_dt(c, local);
}

// ------------------------------------------------------------------------------------------
// Step size of the trial (reference choose_dt_kernel semantics) and the first RK stage.
__device__ __forceinline__ bool choose_dt(const Control* c, double fixed, double* dt_out) {
  double dt = c->retrying ? c->dt : (fixed > 0 ? fixed : __longlong_as_double(static_cast<long long>(c->dt_min)));
  const double remaining = c->target - c->time;
  dt = fmin(dt, remaining);
  const double tol = 32 * DBL_EPSILON * fmax(fmax(fabs(c->target), fabs(c->time)), dt);
  if (remaining - dt <= tol) dt = remaining;
  *dt_out = dt;
  return isfinite(dt) && dt > 0 && c->time + dt != c->time;
}

template <class T>
__global__ __launch_bounds__(kMaxTile) void k_stage0(ArenaPtrs<T> a) {
  const int m = a.cell_block_mesh[blockIdx.x];
  Control* c = a.ctrl + m;
  if (!running(c)) return;
  const MeshDesc<T>& d = a.desc[m];
  const bool leader = blockIdx.x == static_cast<unsigned>(d.first_cell_block) && threadIdx.x == 0;
  if (c->bad & BAD_BASE) {  // the base residual failed: nonphysical base state
    if (leader) c->fatal = 2;
    return;
  }
  double dt;
  if (!choose_dt(c, d.dt_fixed, &dt)) {
    if (leader) c->fatal = 4;
    return;
  }
  if (leader) {
    c->trial_dt = dt;
    if (!c->retrying) c->attempt = 0;
  }
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.cell_begin + d.n_cells) return;
  const int nc = a.nc, cur = c->cur;
  const double* u = a.U[cur];
  double* out = a.U[cur ^ 1];
  for (int v = 0; v < d.nv; ++v) out[v * nc + k] = u[v * nc + k] + dt * static_cast<double>(a.R0[v * nc + k]);
  if (!d.small) {
    T w[4], phi = T(0);
    if (!to_primitive(out, nc, k, d.gamma, d.sa != 0, w, &phi)) flag_bad(c, BAD_STAGE);
    else store_primitive(a, k, w, phi, d.sa != 0);
  }
}

// SRD: q_g = sum_k w_gk u_k, then u_k = mean of q over the groups holding k, then primitives.
// `stage` selects the candidate: 0 and 2 write U[cur ^ 1], 1 writes U2.
template <class T>
__global__ __launch_bounds__(kMaxTile) void k_group(ArenaPtrs<T> a, int stage) {
  const int m = a.cell_block_mesh[blockIdx.x];
  Control* c = a.ctrl + m;
  if (!healthy(c)) return;
  const MeshDesc<T>& d = a.desc[m];
  if (!d.small) return;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= d.cell_begin + d.n_cells) return;
  const int nc = a.nc;
  const double* u = stage == 1 ? a.U2 : a.U[c->cur ^ 1];
  double q0 = 0, q1 = 0, q2 = 0, q3 = 0;
  for (int j = a.grp_off[g]; j < a.grp_off[g + 1]; ++j) {
    const int k = a.grp_mem[j];
    const double w = static_cast<double>(a.grp_w[j]);
    q0 += w * u[k];
    q1 += w * u[nc + k];
    q2 += w * u[2 * nc + k];
    q3 += w * u[3 * nc + k];
  }
  a.QHAT[g] = q0;
  a.QHAT[nc + g] = q1;
  a.QHAT[2 * nc + g] = q2;
  a.QHAT[3 * nc + g] = q3;
}

template <class T>
__global__ __launch_bounds__(kMaxTile) void k_redist(ArenaPtrs<T> a, int stage) {
  const int m = a.cell_block_mesh[blockIdx.x];
  Control* c = a.ctrl + m;
  if (!healthy(c)) return;
  const MeshDesc<T>& d = a.desc[m];
  if (!d.small) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.cell_begin + d.n_cells) return;
  const int nc = a.nc;
  double* u = stage == 1 ? a.U2 : a.U[c->cur ^ 1];
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int j = a.mem_off[k]; j < a.mem_off[k + 1]; ++j) {
    const int g = a.mem_grp[j];
    s0 += a.QHAT[g];
    s1 += a.QHAT[nc + g];
    s2 += a.QHAT[2 * nc + g];
    s3 += a.QHAT[3 * nc + g];
  }
  const double inv = static_cast<double>(a.mem_inv[k]);
  u[k] = s0 * inv;
  u[nc + k] = s1 * inv;
  u[2 * nc + k] = s2 * inv;
  u[3 * nc + k] = s3 * inv;
  T w[4], phi = T(0);
  if (!to_primitive(u, nc, k, d.gamma, d.sa != 0, w, &phi)) flag_bad(c, BAD_STAGE);
  else store_primitive(a, k, w, phi, d.sa != 0);
}

// ------------------------------------------------------------------------------------------
// Start the next trial: finished when the target is reached, fatal 3 at the step cap.
__device__ __forceinline__ void begin_trial(Control* c) {
  c->active = 0;
  if (c->fatal || c->finished) return;
  if (c->time >= c->target) {
    c->finished = 1;
    return;
  }
  if (c->steps >= c->maxsteps) {
    c->fatal = 3;
    return;
  }
  c->bad = 0;
  c->active = 1;
  if (!c->r0_valid) c->dt_min = kInfBits;
}

__global__ void k_commit(Control* ctrl, int n) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  Control* c = ctrl + m;
  if (c->active) {
    if (c->fatal) {
      c->w_valid = 0;
    } else if (c->bad) {  // reject: same base state, half the step
      ++c->rejected;
      ++c->attempt;
      c->retrying = 1;
      c->dt = c->trial_dt * 0.5;
      if (c->attempt >= 21) c->fatal = 5;
      c->w_valid = 0;
      c->r0_valid = 1;
    } else {  // accept: the candidate buffer becomes the state
      const double remaining = c->target - c->time;
      c->time = c->trial_dt == remaining ? c->target : c->time + c->trial_dt;
      ++c->steps;
      c->dt = c->trial_dt;
      c->retrying = 0;
      c->attempt = 0;
      c->cur ^= 1;
      c->w_valid = 1;
      c->r0_valid = 0;
      c->finished = c->time >= c->target;
    }
  }
  begin_trial(c);
}

__global__ void k_target(Control* ctrl, int n, double target, long long maxsteps, const double* targets,
                         const long long* maxsteps_each) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  Control* c = ctrl + m;
  c->target = targets ? targets[m] : target;
  c->maxsteps = maxsteps_each ? maxsteps_each[m] : maxsteps;
  c->finished = c->time >= c->target;
  begin_trial(c);
}

// Sum over the block of several doubles (same wave64 + LDS pattern as block_min); valid in thread 0.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N]) {
  __shared__ double part[N][kMaxTile / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = 0; i < N; ++i) {
    for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off, 64);
    if (lane == 0) part[i][wave] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int i = 0; i < N; ++i) {
      double t = part[i][0];
      for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) t += part[i][w];
      v[i] = t;
    }
}

// Steady monitor, cells: volume-weighted |R0| over the residual region (FP64 accumulation;
// reference src/quick_airfoil.cpp:76-81).
template <class T>
__global__ __launch_bounds__(kMaxTile) void k_monitor_cells(ArenaPtrs<T> a) {
  const int m = a.cell_block_mesh[blockIdx.x];
  const MeshDesc<T>& d = a.desc[m];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (k < d.cell_begin + d.n_cells) {
    const double* f = a.mon_frame + 6 * m;
    const double x0 = f[0], x1 = f[1], y0 = f[2], y1 = f[3];
    const double cx = a.cell_cx[k], cy = a.cell_cy[k];
    if (!(cx < x0 || cx > x1 || cy < y0 || cy > y1)) {
      const double vol = static_cast<double>(a.vol[k]);
      acc[5] = vol;
      for (int v = 0; v < 5; ++v)
        if (v < d.nv) acc[v] = vol * fabs(static_cast<double>(a.R0[v * a.nc + k]));
    }
  }
  block_sum(acc);
  if (threadIdx.x == 0)
    for (int i = 0; i < 6; ++i) a.MON_CELL[6 * blockIdx.x + i] = acc[i];
}

// Steady monitor, walls: pressure force minus viscous shear and the moment about (mx, my)
// (reference src/quick_airfoil.cpp:83-85).
template <class T>
__global__ __launch_bounds__(kMaxTile) void k_monitor_walls(ArenaPtrs<T> a) {
  const int b = blockIdx.x;
  const int m = a.face_block_mesh[a.n_face_blocks + b];
  const MeshDesc<T>& d = a.desc[m];
  const int j = b * blockDim.x + threadIdx.x;
  double acc[3] = {0, 0, 0};
  if (j < d.wall_begin + d.n_walls) {
    const double mx = a.mon_frame[6 * m + 4], my = a.mon_frame[6 * m + 5];
    const int k = a.wall_owner[j];
    const double p = static_cast<double>(a.W[3 * a.nc + k]), len = static_cast<double>(a.wall_len[j]);
    double fx = p * static_cast<double>(a.wall_nx[j]) * len, fy = p * static_cast<double>(a.wall_ny[j]) * len;
    if (d.laminar) {
      fx -= static_cast<double>(a.WFLUX[j]);
      fy -= static_cast<double>(a.WFLUX[a.nw + j]);
    }
    acc[0] = fx;
    acc[1] = fy;
    acc[2] = (a.wall_mx[j] - mx) * fy - (a.wall_my[j] - my) * fx;
  }
  block_sum(acc);
  if (threadIdx.x == 0)
    for (int i = 0; i < 3; ++i) a.MON_WALL[3 * b + i] = acc[i];
}

inline int blocks_of(int n, int tile) { return (n + tile - 1) / tile; }

}  // namespace

#define HIPAMR_LAUNCH(name, kernel, grid, ...)                      \
  do {                                                               \
    if (hook) hook->before(name);                                    \
    hipLaunchKernelGGL(kernel, grid, tile, 0, s, __VA_ARGS__);       \
    if (hook) hook->after(name);                                     \
    ++launches;                                                      \
  } while (0)

template <class T>
int launch_residual(const ArenaPtrs<T>& a, hipStream_t s, LaunchHook* hook) {
  const dim3 tile(a.tile), cells(a.n_cell_blocks), faces(a.n_face_blocks + a.n_wall_blocks);
  int launches = 0;
  HIPAMR_LAUNCH("base_primitive", k_base_prim<T>, cells, a);
  if (a.any_laminar) HIPAMR_LAUNCH("gradient_base", (k_grad<T, PASS_BASE>), cells, a);
  HIPAMR_LAUNCH("face_base", (k_face<T, PASS_BASE>), faces, a);
  HIPAMR_LAUNCH("collect_base", (k_collect<T, 0>), cells, a);
  if (a.any_visc_group) {
    HIPAMR_LAUNCH("diffusion_group", k_diffusion_group<T>, cells, a);
    HIPAMR_LAUNCH("dt_group", k_dt_group<T>, cells, a);
  }
  return launches;
}

template <class T>
int launch_trial(const ArenaPtrs<T>& a, hipStream_t s, LaunchHook* hook) {
  const dim3 tile(a.tile), cells(a.n_cell_blocks), faces(a.n_face_blocks + a.n_wall_blocks);
  int launches = launch_residual(a, s, hook);
  HIPAMR_LAUNCH("stage0", k_stage0<T>, cells, a);
  if (a.any_small) {
    HIPAMR_LAUNCH("srd_group", k_group<T>, cells, a, 0);
    HIPAMR_LAUNCH("srd_redistribute", k_redist<T>, cells, a, 0);
  }
  for (int stage = 1; stage <= 2; ++stage) {
    if (a.any_laminar) HIPAMR_LAUNCH("gradient_stage", (k_grad<T, PASS_STAGE>), cells, a);
    HIPAMR_LAUNCH("face_stage", (k_face<T, PASS_STAGE>), faces, a);
    if (stage == 1) HIPAMR_LAUNCH("collect_rk1", (k_collect<T, 1>), cells, a);
    else HIPAMR_LAUNCH("collect_rk2", (k_collect<T, 2>), cells, a);
    if (a.any_small) {
      HIPAMR_LAUNCH("srd_group", k_group<T>, cells, a, stage);
      HIPAMR_LAUNCH("srd_redistribute", k_redist<T>, cells, a, stage);
    }
  }
  {
    const dim3 tile(256);
    HIPAMR_LAUNCH("commit", k_commit, dim3(blocks_of(a.n_meshes, 256)), a.ctrl, a.n_meshes);
  }
  return launches;
}
#undef HIPAMR_LAUNCH

template <class T>
void launch_monitor(const ArenaPtrs<T>& a, hipStream_t s) {
  hipLaunchKernelGGL(k_monitor_cells<T>, dim3(a.n_cell_blocks), dim3(a.tile), 0, s, a);
  if (a.n_wall_blocks > 0) hipLaunchKernelGGL(k_monitor_walls<T>, dim3(a.n_wall_blocks), dim3(a.tile), 0, s, a);
}

void launch_target(Control* ctrl, int n, double target, long long maxsteps, const double* targets,
                   const long long* maxsteps_each, hipStream_t s) {
  hipLaunchKernelGGL(k_target, dim3(blocks_of(n, 256)), dim3(256), 0, s, ctrl, n, target, maxsteps, targets,
                     maxsteps_each);
}

template int launch_trial<double>(const ArenaPtrs<double>&, hipStream_t, LaunchHook*);
template int launch_trial<float>(const ArenaPtrs<float>&, hipStream_t, LaunchHook*);
template int launch_residual<double>(const ArenaPtrs<double>&, hipStream_t, LaunchHook*);
template int launch_residual<float>(const ArenaPtrs<float>&, hipStream_t, LaunchHook*);
template void launch_monitor<double>(const ArenaPtrs<double>&, hipStream_t);
template void launch_monitor<float>(const ArenaPtrs<float>&, hipStream_t);

}  // namespace hipamr
