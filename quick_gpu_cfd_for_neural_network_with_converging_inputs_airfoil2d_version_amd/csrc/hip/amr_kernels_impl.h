// amr_kernels_impl.h -- the explicit cut-cell SSP-RK3 trial as gfx950 kernels, templated on
// the work precision; amr_kernels_f64.hip and amr_kernels_f32.hip instantiate it (the FP32 unit
// is compiled with reciprocal-based division, the FP64 unit with IEEE division).
//
// One trial (reference launch_trial, src/static_amr2d_cuda.cu:241-268) is, per RHS pass:
//   k_grad   cells: LSQ velocity gradient, SA-neg gradients and point model (reference
//            :71-91, src/static_amr2d_sa_cuda.cuh:2-11) into the gradient vector GREC
//   k_flux   tiles: (A) each face touching the tile once (boundary state, HLLC / AUSM+-up,
//            viscous and SA face flux, reference :92-143, src/static_amr2d_sa_cuda.cuh:12-29)
//            into an LDS slot; (B) per cell the sum of its faces' slots, the no-slip cut-wall
//            shear of its wall fragments and r = -(sum F + p n l - wall shear)/V (reference
//            :144-163); the base pass also forms the local CFL / viscous / SA step and reduces
//            it (wave shuffles + LDS + one 64-bit atomicMin per block), stage passes apply the
//            SSP-RK3 combination (reference :188-194) and write the candidate's primitives,
//            viscosity and eddy viscosity.
// A face shared by two tiles is evaluated by both with the same inputs and operations, so both
// see the same flux bits; no flux array is written or re-read.  Cell data that faces and
// gradient stencils gather lives in short vectors per cell (AoS per purpose: 4 T, 3 T in FP32
// for p / mu / eddy viscosity, 8 T of gradients), so one neighbour costs whole vector loads;
// each flux block stages its own tile's primitive vectors in LDS.  Primitive vectors rotate
// through three buffers: a stage reads one and writes the next state's into another, and the
// base state's survive a rejected trial.
// Plus k_base_prim, k_stage0 (dt choice + first stage, reference :170-191), the SRD pair
// k_group / k_redist (reference :195-207), the SRD-grouped viscous limit (reference :164-176)
// and k_commit (accept / reject / begin, reference :54-59,212-216).
// Every kernel serves every mesh of the arena; a block reads its mesh from a block table and
// skips the whole tile when that mesh is not running (block-uniform branch).
#ifndef HIPAMR_AMR_KERNELS_IMPL_H
#define HIPAMR_AMR_KERNELS_IMPL_H

#include <cfloat>

#include "amr_arena.h"

namespace hipamr {
namespace {

constexpr int kMaxTile = 256;
constexpr unsigned long long kInfBits = 0x7ff0000000000000ull;

enum : int { PASS_BASE = 0, PASS_STAGE = 1 };

__device__ __forceinline__ bool running(const Control* c) { return c->active && !c->fatal; }
__device__ __forceinline__ bool healthy(const Control* c) { return running(c) && !c->bad; }
// Whether a pass runs for a slot.  `active`, `fatal` and `r0_valid` change between the launches
// of a pass's kernels only (every kernel that uses barriers), while `bad` is raised by other
// blocks during a launch: pass_live leaves it out, and a kernel whose waves meet at barriers
// takes `bad` from one snapshot per block (taken by thread 0 before its first barrier), so its
// waves return or go on together and never read LDS that other waves left unstaged.
__device__ __forceinline__ bool pass_live(const Control* c, int pass) {
  return running(c) && (pass == PASS_STAGE || !c->r0_valid);
}
__device__ __forceinline__ bool pass_on(const Control* c, int pass) { return pass_live(c, pass) && !c->bad; }
// Admissibility bits: 1 = the base pass failed (fatal 2 at the dt choice), 2 = a candidate
// stage failed (the trial is rejected).  Separate bits keep the dt choice race-free.
enum : int { BAD_BASE = 1, BAD_STAGE = 2 };
__device__ __forceinline__ void flag_bad(Control* c, int bit) { atomicOr(&c->bad, bit); }

// Three primitive buffers per cell.  A = wbuf holds the primitives of U[cur]; stage 0 writes
// B = A + 1 (mod 3), read by stage 1, which writes C = A + 2, read by stage 2, which writes the
// candidate state's primitives into B.  An accepted trial makes B the new A; a rejected one
// leaves A intact, so the base primitives are recomputed only after the host changes a state.
__device__ __forceinline__ int read_buffer(const Control* c, int stage) { return (c->wbuf + stage) % 3; }
__device__ __forceinline__ int write_buffer(const Control* c, int stage) { return (c->wbuf + (stage == 1 ? 2 : 1)) % 3; }

// FP32 work precision, mesh without SRD: the RK stage states U1, U2 are kept as T increments
// over the FP64 state (U1 = U + D1, U2 = U + D2), half the bytes of FP64 copies.  The
// increments are O(dt r), so their FP32 rounding is far below the FP32 flux error; the state and
// the final combination stay FP64.  FP64 work precision and SRD meshes use FP64 stage states.
// Decided per mesh, so a slot computes the same bits whatever else shares its arena.
template <class T>
__device__ __forceinline__ bool stage_deltas(const MeshDesc<T>& d) {
  return sizeof(T) == 4 && !d.small;
}

// Cell tile of this block and its slot, or {-1, -1}: one 8-byte load, so that the slot's control
// record is one dependent load away from the kernel's start.  The dispatcher hands consecutive
// blocks to the 8 XCDs in turn, each with its own L2.  The host-built table splits every slot's
// tiles into 8 contiguous runs, one per XCD: neighbouring tiles share an L2, and whichever slots
// are running, their work is spread over all XCDs.
template <class T>
__device__ __forceinline__ int2 cell_block(const ArenaPtrs<T>& a) {
  return a.block_rec[blockIdx.x];
}

// ---- cell records ---------------------------------------------------------------------------
template <class T>
struct CellPrim {
  T rho, u, v, p, phi, mu, eddy;
};
template <class T>
struct CellGrad {
  T ux, uy, vx, vy, sx, sy, sdiff;
};

template <class T>
struct CellVel {
  T rho, u, v, phi;
};

// The last read of a per-cell plane in a trial (the state in the last stage, the residual and
// the increment where they are consumed): a non-temporal load, so that data no later kernel
// reads does not displace what the next kernels gather.  (Non-temporal stores of the records and
// planes that the next kernel reads were measured and dropped: they bypass the device's last-level
// cache, and the kernel after them got slower, stage0 49 -> 56 us on 256 LF2 meshes.)
template <class X>
__device__ __forceinline__ X ld_last(const X* p) {
  return __builtin_nontemporal_load(p);
}

// 16-byte vector loads and stores of 4 T (float4, or two double2).
__device__ __forceinline__ void load4(const float* p, float& a, float& b, float& c, float& d) {
  const float4 v = *reinterpret_cast<const float4*>(p);
  a = v.x, b = v.y, c = v.z, d = v.w;
}
__device__ __forceinline__ void load4(const double* p, double& a, double& b, double& c, double& d) {
  const double2 v = *reinterpret_cast<const double2*>(p), w = *reinterpret_cast<const double2*>(p + 2);
  a = v.x, b = v.y, c = w.x, d = w.y;
}
// PREC records: 12-byte loads and stores in FP32, the first three of 4 padded words in FP64.
__device__ __forceinline__ void load3(const float* p, float& a, float& b, float& c) {
  const float3 v = *reinterpret_cast<const float3*>(p);
  a = v.x, b = v.y, c = v.z;
}
__device__ __forceinline__ void load3(const double* p, double& a, double& b, double& c) {
  const double2 v = *reinterpret_cast<const double2*>(p);
  a = v.x, b = v.y, c = p[2];
}
__device__ __forceinline__ void store3(float* p, float a, float b, float c) {
  *reinterpret_cast<float3*>(p) = make_float3(a, b, c);
}
__device__ __forceinline__ void store3(double* p, double a, double b, double c) {
  *reinterpret_cast<double2*>(p) = make_double2(a, b);
  p[2] = c;
  p[3] = 0.0;
}
__device__ __forceinline__ void store4(float* p, float a, float b, float c, float d) {
  *reinterpret_cast<float4*>(p) = make_float4(a, b, c, d);
}
__device__ __forceinline__ void store4(double* p, double a, double b, double c, double d) {
  *reinterpret_cast<double2*>(p) = make_double2(a, b);
  *reinterpret_cast<double2*>(p + 2) = make_double2(c, d);
}

template <class T>
__device__ __forceinline__ CellVel<T> load_vel(const ArenaPtrs<T>& a, int buf, int k) {
  CellVel<T> v;
  load4(a.VREC[buf] + 4 * k, v.rho, v.u, v.v, v.phi);
  return v;
}
template <class T>
__device__ __forceinline__ CellPrim<T> load_prim(const ArenaPtrs<T>& a, int buf, int k) {
  CellPrim<T> w;
  load4(a.VREC[buf] + 4 * k, w.rho, w.u, w.v, w.phi);
  load3(a.PREC[buf] + prec_width<T>() * k, w.p, w.mu, w.eddy);
  return w;
}
template <class T>
__device__ __forceinline__ CellGrad<T> load_grad(const ArenaPtrs<T>& a, int k) {
  CellGrad<T> g;
  T pad;
  load4(a.GREC + 8 * k, g.ux, g.uy, g.vx, g.vy);
  load4(a.GREC + 8 * k + 4, g.sx, g.sy, g.sdiff, pad);
  return g;
}
template <class T>
__device__ __forceinline__ void store_prim(const ArenaPtrs<T>& a, int buf, int k, const CellPrim<T>& w) {
  store4(a.VREC[buf] + 4 * k, w.rho, w.u, w.v, w.phi);
  store3(a.PREC[buf] + prec_width<T>() * k, w.p, w.mu, w.eddy);
}
template <class T>
// The gradient vector's 8th slot carries the cell's SA-neg source (read by the cell itself only).
__device__ __forceinline__ void store_grad(const ArenaPtrs<T>& a, int k, const CellGrad<T>& g, T sa_source) {
  store4(a.GREC + 8 * k, g.ux, g.uy, g.vx, g.vy);
  store4(a.GREC + 8 * k + 4, g.sx, g.sy, g.sdiff, sa_source);
}

// Primitive vectors of the block's own tile staged in LDS: a neighbour inside the tile (most of
// them, with cells in Hilbert order) is read from LDS instead of being gathered from L2 / HBM
// again.  The flux kernels also keep one chunk of face results there (FaceSlots).  Both live in
// the launch's dynamic LDS, sized by the tile, so small tiles keep their occupancy.  Gradient
// records are not staged: the LDS they took (32 B per cell) is worth more as occupancy (the
// flux kernels run at 6 waves per SIMD, bounded by registers and LDS together).
extern __shared__ __align__(16) unsigned char tile_lds[];

template <class T>
struct TileCells {
  T* v;  // 4 per cell: rho, u, v, phi
  T* p;  // 4 per cell: p, mu, eddy, 0
  __device__ explicit TileCells(int tile) : v(reinterpret_cast<T*>(tile_lds)), p(v + 4 * tile) {}
};

// Results of one chunk of face items: f = mass, x- and y-momentum, energy flux (times the
// length, convective minus viscous); stage passes: sf = SA flux; base pass: x = spectral radius
// x length, viscous and SA dt coefficients and the SA flux (4 per item, so that a base item takes
// 8 words).
template <class T>
struct FaceSlots {
  T* f;
  T* sf;  // stage passes
  T* x;   // base pass
  __device__ FaceSlots(int tile, bool base)
      : f(reinterpret_cast<T*>(tile_lds) + 8 * tile), sf(f + 4 * item_chunk(tile, base)), x(sf) {}
};

template <class T>
__device__ __forceinline__ void stage_cell(const TileCells<T>& t, int i, const CellPrim<T>& w) {
  store4(t.v + 4 * i, w.rho, w.u, w.v, w.phi);
  store4(t.p + 4 * i, w.p, w.mu, w.eddy, T(0));
}
template <class T>
__device__ __forceinline__ CellPrim<T> tile_prim(const TileCells<T>& t, int i) {
  CellPrim<T> w;
  T pad;
  load4(t.v + 4 * i, w.rho, w.u, w.v, w.phi);
  load4(t.p + 4 * i, w.p, w.mu, w.eddy, pad);
  return w;
}

// Face geometry record: length, dx, dy, ex, ey, distance.
template <class T>
struct FaceGeo {
  T len, dx, dy, ex, ey, dist;
};

// Face item i of a tile: 4 T words of ITEM_A (two ints: left cell or -1; right cell + 1 in bits
// 0..kItemRightBits-1, the face info above, bit 31 set for irregular geometry; then length and
// d_n), and for an irregular item 4 T words of ITEM_B (d_t, ex, ey, distance).  A regular item's
// geometry is d_t = ex = ey = 0, distance = |d_n|: the same bits the host would have stored.
template <class T>
struct FaceItem {
  int left, right, info;
  FaceGeo<T> geo;
};
template <class T>
__device__ __forceinline__ FaceItem<T> finish_item(int w0, int w1, T len, T dn, const T* b) {
  FaceItem<T> it;
  it.left = w0;
  it.right = (w1 & ((1 << kItemRightBits) - 1)) - 1;
  it.info = (w1 >> kItemRightBits) & kItemInfoMask;
  T dt = T(0), ex = T(0), ey = T(0), dist = fabs(dn);
  if (w1 < 0) load4(b, dt, ex, ey, dist);
  const bool xaxis = (it.info & 1) == 0;
  it.geo = FaceGeo<T>{len, xaxis ? dn : dt, xaxis ? dt : dn, ex, ey, dist};
  return it;
}
__device__ __forceinline__ FaceItem<float> load_item(const float* A, const float* B, size_t i) {
  const int4 a = *reinterpret_cast<const int4*>(A + 4 * i);
  return finish_item<float>(a.x, a.y, __int_as_float(a.z), __int_as_float(a.w), B + 4 * i);
}
__device__ __forceinline__ FaceItem<double> load_item(const double* A, const double* B, size_t i) {
  const int4 a = *reinterpret_cast<const int4*>(A + 4 * i);
  const double2 g = *reinterpret_cast<const double2*>(A + 4 * i + 2);
  return finish_item<double>(a.x, a.y, g.x, g.y, B + 4 * i);
}
// The ITEM_A words of an item as loaded (so that the next item's can be in flight while the
// current one is evaluated), and the item they make.
template <class T>
struct ItemWords;
template <>
struct ItemWords<float> {
  int4 a;
};
template <>
struct ItemWords<double> {
  int4 a;
  double2 g;
};
__device__ __forceinline__ ItemWords<float> load_item_words(const float* A, size_t i) {
  return ItemWords<float>{*reinterpret_cast<const int4*>(A + 4 * i)};
}
__device__ __forceinline__ ItemWords<double> load_item_words(const double* A, size_t i) {
  return ItemWords<double>{*reinterpret_cast<const int4*>(A + 4 * i), *reinterpret_cast<const double2*>(A + 4 * i + 2)};
}
__device__ __forceinline__ FaceItem<float> item_of(const ItemWords<float>& w, const float* B, size_t i) {
  return finish_item<float>(w.a.x, w.a.y, __int_as_float(w.a.z), __int_as_float(w.a.w), B + 4 * i);
}
__device__ __forceinline__ FaceItem<double> item_of(const ItemWords<double>& w, const double* B, size_t i) {
  return finish_item<double>(w.a.x, w.a.y, w.g.x, w.g.y, B + 4 * i);
}

// U -> cell primitives (computed in FP64 from the FP64 state, stored as T), plus molecular
// and eddy viscosity for viscous meshes; false if not admissible (reference primitive rules,
// src/static_amr2d.cpp:143-154, and a valid viscosity).
// q: rho, rho u, rho v, energy, rho nutilde (the last only with SA).
template <class T>
__device__ __forceinline__ bool make_prim_q(const double (&q)[5], const MeshDesc<T>& d, CellPrim<T>& w) {
  const double rho = q[0], mx = q[1], my = q[2], energy = q[3];
  const double ir = 1.0 / rho;
  const double vx = mx * ir, vy = my * ir;
  const double p = (static_cast<double>(d.gamma) - 1.0) * (energy - 0.5 * (mx * vx + my * vy));
  bool ok = isfinite(rho) && rho > 0 && isfinite(mx) && isfinite(my) && isfinite(energy) && isfinite(p) && p > 0;
  w.rho = static_cast<T>(rho);
  w.u = static_cast<T>(vx);
  w.v = static_cast<T>(vy);
  w.p = static_cast<T>(p);
  w.phi = T(0);
  if (d.sa) {
    const double ph = q[4] * ir;
    w.phi = static_cast<T>(ph);
    ok = ok && isfinite(ph);
  }
  ok = ok && w.rho > T(0) && w.p > T(0) && isfinite(w.rho) && isfinite(w.p);
  w.mu = T(0);
  w.eddy = T(0);
  if (d.laminar && ok) {
    const T mu = viscosity(d.transport, w.p / (w.rho * d.transport.gas_constant));
    ok = mu > T(0) && isfinite(mu);
    w.mu = mu;
    if (d.sa) w.eddy = w.rho * sa_eddy(w.phi, mu / w.rho);
  }
  return ok;
}
template <class T>
__device__ __forceinline__ bool make_prim(const double* u, int nc, int k, const MeshDesc<T>& d, CellPrim<T>& w) {
  const double q[5] = {u[k], u[nc + k], u[2 * nc + k], u[3 * nc + k], d.sa ? u[4 * nc + k] : 0.0};
  return make_prim_q(q, d, w);
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
  const int2 blk = cell_block(a);
  const int tile = blk.x;
  if (tile < 0) return;
  const int m = blk.y;
  Control* c = a.ctrl + m;
  if (!pass_on(c, PASS_BASE) || c->w_valid) return;
  const MeshDesc<T>& d = a.desc[m];
  const int k = tile * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
  if (k >= d.cell_begin + d.n_cells) return;
  CellPrim<T> w;
  if (!make_prim(a.U[c->cur], a.nc, k, d, w)) {
    flag_bad(c, BAD_BASE);
    atomicExch(&c->fatal, 2);  // the accepted state itself is not admissible
    return;
  }
  store_prim(a, c->wbuf, k, w);
}

// ------------------------------------------------------------------------------------------
// Step size of the trial (reference choose_dt_kernel semantics, src/static_amr2d_cuda.cu:177-187).
__device__ __forceinline__ bool choose_dt(const Control* c, double fixed, double* dt_out) {
  double dt = c->retrying ? c->dt : (fixed > 0 ? fixed : __longlong_as_double(static_cast<long long>(c->dt_min)));
  const double remaining = c->target - c->time;
  dt = fmin(dt, remaining);
  const double tol = 32 * DBL_EPSILON * fmax(fmax(fabs(c->target), fabs(c->time)), dt);
  if (remaining - dt <= tol) dt = remaining;
  *dt_out = dt;
  return isfinite(dt) && dt > 0 && c->time + dt != c->time;
}

// The dt choice that opens the first RK stage, evaluated by every block of the slot from the
// same control fields; the slot's first block records the trial step or the fatal code (2: the
// base residual failed, 4: dt underflow).  False: no stage this trial.
template <class T>
__device__ __forceinline__ bool open_stage0(Control* c, const MeshDesc<T>& d, int tile, double* dt) {
  const bool leader = tile == d.first_cell_block && threadIdx.x == 0;
  if (c->bad & BAD_BASE) {  // the base residual failed: nonphysical base state
    if (leader) c->fatal = 2;
    return false;
  }
  if (!choose_dt(c, d.dt_fixed, dt)) {
    if (leader) c->fatal = 4;
    return false;
  }
  if (leader) {
    c->trial_dt = *dt;
    if (!c->retrying) c->attempt = 0;
  }
  return true;
}

// First SSP-RK3 stage of cell k, U1 = U + dt R0, as conserved values q (FP64).  With FP32
// increments (stage_deltas) the increment is rounded to T first, D1 = T(dt R0), exactly as the
// second stage re-forms it; otherwise U1 is also stored in U[cur ^ 1] when `store`.
template <class T>
__device__ __forceinline__ void first_stage(const ArenaPtrs<T>& a, const MeshDesc<T>& d, const Control* c, int k,
                                            double dt, bool store, double (&q)[5]) {
  const int nc = a.nc, cur = c->cur;
  const double* u = a.U[cur];
  for (int v = 0; v < 5; ++v) q[v] = 0.0;
  if (stage_deltas(d)) {
    // Every load of the cell first (the fifth variable only with SA), then the sums: the loads of
    // a loop over the run-time variable count would each wait for the one before.
    double uv[5] = {0, 0, 0, 0, 0};
    T rv[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      uv[v] = u[v * nc + k];
      rv[v] = a.R0[v * nc + k];
    }
    if (d.nv == 5) {
      uv[4] = u[4 * nc + k];
      rv[4] = a.R0[4 * nc + k];
    }
#pragma unroll
    for (int v = 0; v < 5; ++v)
      if (v < d.nv) q[v] = uv[v] + static_cast<double>(static_cast<T>(dt * static_cast<double>(rv[v])));
  } else {
    double uv[5] = {0, 0, 0, 0, 0};
    T rv[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      uv[v] = u[v * nc + k];
      rv[v] = a.R0[v * nc + k];
    }
    if (d.nv == 5) {
      uv[4] = u[4 * nc + k];
      rv[4] = a.R0[4 * nc + k];
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      if (v < d.nv) {
        q[v] = uv[v] + dt * static_cast<double>(rv[v]);
        if (store) a.U[cur ^ 1][v * nc + k] = q[v];
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// LSQ gradients of the stage's state (and the SA-neg point source) into the gradient vector
// GREC.  STAGE 0 = base pass, 1 and 2 = RK stages.
template <class T, int STAGE>
__global__ __launch_bounds__(kMaxTile) void k_grad(ArenaPtrs<T> a) {
  const int2 blk = cell_block(a);
  const int tile = blk.x;
  if (tile < 0) return;
  const int m = blk.y;
  Control* c = a.ctrl + m;
  if (!pass_live(c, STAGE == 0 ? PASS_BASE : PASS_STAGE)) return;
  const MeshDesc<T>& d = a.desc[m];
  if (!d.laminar) return;
  __shared__ int failed;  // block snapshot of the slot's admissibility bits
  if (threadIdx.x == 0) failed = c->bad;
  const int first = tile * static_cast<int>(blockDim.x);
  const int k = first + static_cast<int>(threadIdx.x);
  const int staged = min(static_cast<int>(blockDim.x), d.cell_begin + d.n_cells - first);  // valid cells of the tile
  const int rb = read_buffer(c, STAGE);
  const bool sa = d.sa != 0;
  T* const vel = reinterpret_cast<T*>(tile_lds);  // the tile's own velocity vectors (4 per cell)
  CellPrim<T> own{};
  int begin = 0, end = 0;  // the cell's terms, loaded with its record (not after the barrier)
  T wall_distance = T(1);
  if (k < d.cell_begin + d.n_cells) {
    own = load_prim(a, rb, k);
    begin = a.term_off[k];
    end = a.term_off[k + 1];
    if (sa) wall_distance = a.wall_dist[k];
    store4(vel + 4 * threadIdx.x, own.rho, own.u, own.v, own.phi);
  }
  __syncthreads();
  if (failed || k >= d.cell_begin + d.n_cells) return;
  const T rho = own.rho, u = own.u, v = own.v, phi = own.phi;
  T g0 = 0, g1 = 0, g2 = 0, g3 = 0, sx = 0, sy = 0, rx = 0, ry = 0;
  // The terms in batches of kTerms (nine cells in ten have four): every load of a batch (term
  // records, then the neighbours' vectors) is issued before the first sum, so a cell waits for
  // two memory round trips per batch instead of two per term.  The sums run term by term in the
  // stored order, as one term at a time would.
  constexpr int kTerms = 4;
  for (int j0 = begin; j0 < end; j0 += kTerms) {
    int src[kTerms];
    T wx[kTerms], wy[kTerms];
#pragma unroll
    for (int q = 0; q < kTerms; ++q) {
      const int j = min(j0 + q, end - 1);  // terms past the end are loaded in range and not summed
      const TermRec<T> t = a.terms[j];
      src[q] = t.src;
      wx[q] = t.wx;
      wy[q] = t.wy;
    }
    // Neighbour vectors without branches: one generic-address load per term, from the tile's LDS
    // copy when the neighbour is in the tile, else from the record array (a side term reads the
    // cell's own LDS copy and discards it), so the batch's loads are all in flight together.
    CellVel<T> nb[kTerms];
#pragma unroll
    for (int q = 0; q < kTerms; ++q) {
      const int local = src[q] - first;
      const T* from = src[q] < 0 ? vel + 4 * threadIdx.x
                                 : (local >= 0 && local < staged ? vel + 4 * local : a.VREC[rb] + 4 * static_cast<size_t>(src[q]));
      load4(from, nb[q].rho, nb[q].u, nb[q].v, nb[q].phi);
    }
    // Neighbour (u, v, phi, rho) of each term: 0, 0, 0 and the cell's own density for a side
    // without a wall velocity (-1), the side's wall velocity for a wall side (<= -2).
    T uu[kTerms], vv[kTerms], ph[kTerms], rr[kTerms];
#pragma unroll
    for (int q = 0; q < kTerms; ++q) {
      const bool cell = src[q] >= 0;
      uu[q] = cell ? nb[q].u : T(0);
      vv[q] = cell ? nb[q].v : T(0);
      ph[q] = cell && sa ? nb[q].phi : T(0);
      rr[q] = cell && sa ? nb[q].rho : rho;
      if (src[q] <= -2) {
        const Side<T>& s = d.side[-2 - src[q]];
        uu[q] = s.wall_velocity[0];
        vv[q] = s.wall_velocity[1];
      }
    }
#pragma unroll
    for (int q = 0; q < kTerms; ++q) {
      if (j0 + q < end) {
        const T du = uu[q] - u, dv = vv[q] - v;
        g0 += wx[q] * du;
        g1 += wy[q] * du;
        g2 += wx[q] * dv;
        g3 += wy[q] * dv;
        if (sa) {
          sx += wx[q] * (ph[q] - phi);
          sy += wy[q] * (ph[q] - phi);
          rx += wx[q] * (rr[q] - rho);
          ry += wy[q] * (rr[q] - rho);
        }
      }
    }
  }
  CellGrad<T> g{g0, g1, g2, g3, sx, sy, T(0)};
  T sa_source = T(0);
  if (sa) {
    const SaPoint<T> pt = sa_point(phi, own.mu / rho, wall_distance, fabs(g2 - g1));
    const T source = rho * (pt.source + T(SaConst<T>::cb2 / SaConst<T>::sigma) * (sx * sx + sy * sy)) -
                     pt.diffusion * (rx * sx + ry * sy);
    g.sdiff = pt.diffusion;
    sa_source = source;
    if (STAGE == 0) a.SRATE[k] = pt.source_rate;  // the local dt of the base pass only
    if (!(pt.diffusion > T(0)) || !isfinite(source) || !isfinite(pt.source_rate))
      flag_bad(c, STAGE == 0 ? BAD_BASE : BAD_STAGE);
  }
  store_grad(a, k, g, sa_source);
}

// ------------------------------------------------------------------------------------------
// One face seen from one of its cells: boundary state, inviscid flux in the face frame, viscous
// and SA face fluxes.  Everything stays in registers: the face-normal component and the two
// sides are chosen with selects, never by indexing a local array or selecting between the
// addresses of two records (either would put them in scratch memory).
template <class T>
struct FaceFlux {
  T f0, fx, fy, f3, sflux, speed, fcoef, sfcoef;
  bool ok;
};

// Field-wise choice between two register records (no address of either is taken).
template <class T>
__device__ __forceinline__ CellPrim<T> pick(bool first, const CellPrim<T>& a, const CellPrim<T>& b) {
  return CellPrim<T>{first ? a.rho : b.rho, first ? a.u : b.u,     first ? a.v : b.v,      first ? a.p : b.p,
                     first ? a.phi : b.phi, first ? a.mu : b.mu, first ? a.eddy : b.eddy};
}
template <class T>
__device__ __forceinline__ CellGrad<T> pick(bool first, const CellGrad<T>& a, const CellGrad<T>& b) {
  return CellGrad<T>{first ? a.ux : b.ux, first ? a.uy : b.uy, first ? a.vx : b.vx,         first ? a.vy : b.vy,
                     first ? a.sx : b.sx, first ? a.sy : b.sy, first ? a.sdiff : b.sdiff};
}

// L, R (and Lg, Rg) are the records of the face's left and right cells; on a boundary face both
// are the inside cell's.  `in`, `ing`: the evaluating cell, the inside cell of a boundary face.
template <class T, bool BASE>
__device__ __forceinline__ FaceFlux<T> face_flux(const MeshDesc<T>& d, int4 fi, const FaceGeo<T>& geo,
                                                 const CellPrim<T>& in, const CellGrad<T>& ing, const CellPrim<T>& L,
                                                 const CellPrim<T>& R, const CellGrad<T>& Lg, const CellGrad<T>& Rg) {
  FaceFlux<T> o{};
  o.ok = true;
  const int l = fi.x, r = fi.y, info = fi.z;
  const int axis = info & 1, side = (info >> 1) - 1;
  const bool xaxis = axis == 0;
  const T len = geo.len;
  Prim<T> wl{L.rho, L.u, L.v, L.p};
  Prim<T> wr{R.rho, R.u, R.v, R.p};
  int kind = -1;
  if (side >= 0) {
    // One call site for both orientations: the ghost state goes to the side without a cell.
    const Prim<T> inside{in.rho, in.u, in.v, in.p};
    Prim<T> ghost = inside;
    kind = d.side[side].kind;
    if (!side_state(d.side[side], d.gamma, side, inside, ghost)) {
      o.ok = false;
      return o;
    }
    const bool ghost_left = l < 0;  // field-wise selects keep both states in registers
    wl = Prim<T>{ghost_left ? ghost.r : inside.r, ghost_left ? ghost.u : inside.u, ghost_left ? ghost.v : inside.v,
                 ghost_left ? ghost.p : inside.p};
    wr = Prim<T>{ghost_left ? inside.r : ghost.r, ghost_left ? inside.u : ghost.u, ghost_left ? inside.v : ghost.v,
                 ghost_left ? inside.p : ghost.p};
  }
  // Face frame: n = normal velocity, t = tangential velocity.
  const T nL = xaxis ? wl.u : wl.v, tL = xaxis ? wl.v : wl.u;
  const T nR = xaxis ? wr.u : wr.v, tR = xaxis ? wr.v : wr.u;
  T f[4];
  if (kind == KIND_NOSLIP) {
    f[0] = T(0);
    f[1] = in.p;
    f[2] = T(0);
    f[3] = T(0);
  } else if (d.flux == 0) {
    hllc_flux(wl.r, nL, tL, wl.p, wr.r, nR, tR, wr.p, d.gamma, f);
  } else {
    ausm_flux(wl.r, nL, tL, wl.p, wr.r, nR, tR, wr.p, d.gamma, d.mco, f);
  }
  if (!isfinite(f[0]) || !isfinite(f[1]) || !isfinite(f[2]) || !isfinite(f[3])) {
    o.ok = false;
    return o;
  }
  T F0 = f[0] * len, Fx = (xaxis ? f[1] : f[2]) * len, Fy = (xaxis ? f[2] : f[1]) * len, F3 = f[3] * len;
  if (d.laminar) {
    // Viscous face flux (reference viscous_face, src/static_amr2d_cuda.cu:92-107).  For an
    // interior face the inside cell is the left one.
    // On an interior face every quantity taken from the inside cell below is replaced.
    const bool sa = d.sa != 0;
    const bool interior = side < 0;
    T g0 = ing.ux, g1 = ing.uy, g2 = ing.vx, g3 = ing.vy;
    const T rho_k = in.rho, u_k = in.u, v_k = in.v;
    const T eddy_k = sa ? in.eddy : T(0);
    T mu = in.mu + eddy_k, uf = u_k, vf = v_k;
    T nu = mu / rho_k;
    const T dist = geo.dist;
    if (interior) {
      const T gl[4] = {Lg.ux, Lg.uy, Lg.vx, Lg.vy};
      const T gr[4] = {Rg.ux, Rg.uy, Rg.vx, Rg.vy};
      T g[4];
      corrected_gradient(gl, gr, geo.dx, geo.dy, wr.u - wl.u, wr.v - wl.v, g);
      g0 = g[0];
      g1 = g[1];
      g2 = g[2];
      g3 = g[3];
      uf = T(0.5) * (wl.u + wr.u) + g0 * geo.ex + g1 * geo.ey;
      vf = T(0.5) * (wl.v + wr.v) + g2 * geo.ex + g3 * geo.ey;
      const T temp = T(0.5) * (wl.p / wl.r + wr.p / wr.r) / d.transport.gas_constant;
      const T el = sa ? L.eddy : T(0), er = sa ? R.eddy : T(0);
      mu = viscosity(d.transport, temp) + T(0.5) * (el + er);
      nu = fmax((L.mu + el) / wl.r, (R.mu + er) / wr.r);
    } else {
      const T sign = (side & 1) ? T(1) : T(-1);
      if (kind == KIND_NOSLIP) {
        mu = in.mu;
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
        uf += g0 * geo.ex + g1 * geo.ey;
        vf += g2 * geo.ex + g3 * geo.ey;
      }
    }
    const T g[4] = {g0, g1, g2, g3};
    T fv[4];
    stress_flux(mu, g, xaxis ? T(1) : T(0), xaxis ? T(0) : T(1), uf, vf, fv);
    if (!isfinite(fv[1]) || !isfinite(fv[2]) || !isfinite(fv[3])) o.ok = false;
    Fx -= fv[1] * len;
    Fy -= fv[2] * len;
    F3 -= fv[3] * len;
    if (BASE) o.fcoef = T(4.0 / 3.0) * nu / dist;
    if (sa) {
      // SA face flux: upwind nutilde times the mass flux, minus corrected-gradient diffusion
      // (reference sa_face_wall_point, src/static_amr2d_sa_cuda.cuh:16-25).
      const T mass = F0;
      const T pl = l >= 0 ? L.phi : d.transport.sa_nu_inf;
      const T pr = r >= 0 ? R.phi : d.transport.sa_nu_inf;
      T diff = T(0), fc = T(0);
      if (interior) {
        const T mx = T(0.5) * (Lg.sx + Rg.sx), my = T(0.5) * (Lg.sy + Rg.sy);
        const T corr = ((pr - pl) - mx * geo.dx - my * geo.dy) / (geo.dx * geo.dx + geo.dy * geo.dy);
        const T gn = xaxis ? mx + corr * geo.dx : my + corr * geo.dy;
        const T coeff = T(0.5) * (wl.r * Lg.sdiff + wr.r * Rg.sdiff);
        diff = coeff * gn * len;
        fc = coeff / dist * len;
      } else {
        const bool inflow = l < 0 ? mass > T(0) : mass < T(0);
        const T sign = (side & 1) ? T(1) : T(-1);
        if (kind == KIND_NOSLIP || inflow) {
          const T wallv = kind == KIND_NOSLIP ? T(0) : d.transport.sa_nu_inf;
          const T coeff = kind == KIND_NOSLIP ? in.mu * T(1.0 / SaConst<T>::sigma) : rho_k * ing.sdiff;
          diff = sign * coeff * (wallv - in.phi) / dist * len;
          fc = coeff / dist * len;
        }
      }
      o.sflux = mass * (mass >= T(0) ? pl : pr) - diff;
      if (BASE) o.sfcoef = fc;
      if (!isfinite(o.sflux)) o.ok = false;
    }
  }
  o.f0 = F0;
  o.fx = Fx;
  o.fy = Fy;
  o.f3 = F3;
  if (BASE) o.speed = fmax(fabs(nL) + sqrt(d.gamma * wl.p / wl.r), fabs(nR) + sqrt(d.gamma * wr.p / wr.r));
  return o;
}

// No-slip shear on a cut-wall fragment from its donor cell (reference viscous_wall_point,
// src/static_amr2d_cuda.cu:108-113): fv[1], fv[2] per unit length.
template <class T>
__device__ __forceinline__ void wall_shear(const CellPrim<T>& donor, T nx, T ny, T dd, T* fv) {
  const T du = -donor.u / dd, dv = -donor.v / dd;
  const T g[4] = {nx * du, ny * du, nx * dv, ny * dv};
  stress_flux(donor.mu, g, nx, ny, T(0), T(0), fv);
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

// Accept or reject the finished trial of a slot (reference commit_control_kernel,
// src/static_amr2d_cuda.cu:212-216) and start the next; `bad`: the candidate's admissibility.
__device__ __forceinline__ void commit_trial(Control* c, bool bad) {
  if (c->active) {
    if (c->fatal) {
      c->w_valid = 0;
    } else if (bad) {  // reject: same base state, half the step
      ++c->rejected;
      ++c->attempt;
      c->retrying = 1;
      c->dt = c->trial_dt * 0.5;
      if (c->attempt >= 21) c->fatal = 5;
      c->r0_valid = 1;  // R0 and the base primitives (buffer wbuf) still describe U[cur]
    } else {  // accept: the candidate buffers become the state
      const double remaining = c->target - c->time;
      c->time = c->trial_dt == remaining ? c->target : c->time + c->trial_dt;
      ++c->steps;
      c->dt = c->trial_dt;
      c->retrying = 0;
      c->attempt = 0;
      c->cur ^= 1;
      c->wbuf = (c->wbuf + 1) % 3;
      c->w_valid = 1;
      c->r0_valid = 0;
      c->finished = c->time >= c->target;
    }
  }
  begin_trial(c);
}

// A block of the last stage pass has finished (`failed`: some thread of it found the candidate
// inadmissible).  One 64-bit atomic per block counts the blocks (low word) and the failed ones
// (high word); the block that completes the count commits the trial and clears the count.  Every
// other block has made its last access to the slot's data and control by then, and the flags
// raised in earlier launches are visible, so the commit needs no memory fence.
__device__ __forceinline__ void arrive_and_commit(Control* c, int n_blocks, bool failed) {
  const unsigned long long add = 1ull | (failed ? (1ull << 32) : 0ull);
  const unsigned long long before = atomicAdd(&c->arrivals, add);
  if (static_cast<int>(before & 0xffffffffull) + 1 != n_blocks) return;
  const bool any_failed = (before >> 32) != 0 || failed;
  c->arrivals = 0;
  commit_trial(c, c->bad != 0 || any_failed);
}

// ------------------------------------------------------------------------------------------
// Residual of every cell from its faces and wall fragments.  STAGE 0: base residual R0 and the
// local dt.  1, 2: second and third SSP-RK3 combinations, then the candidate's primitives.
// FP32 register budget: 5 waves per SIMD in the base pass (89 VGPRs, no spill; its LDS admits
// 11 blocks of two waves per CU, 5.5 waves per SIMD, so a 6-wave bound, which spilled, bought no
// occupancy: base flux 151 -> 141 us with the batched entries below) and 7 in the stage passes
// (72 VGPRs, 4 and 6 spilled; LDS for 14 blocks per CU).
// On 256 LF2 meshes: base flux 187 -> 177 us, stage flux 188 -> 181 us against the unbounded
// kernels; 8 waves (64 VGPRs) spills 11-14 and is slower.  FP64 (190-200 VGPRs) has no bound.
// SPREAD threads per cell (1 or 2): a tile's face items are spread over SPREAD * tile threads,
// the cell work stays on the first tile of them (a separate instantiation, so that the one-thread
// kernel keeps its registers).
template <class T, int STAGE, int SPREAD>
__global__ __launch_bounds__(kMaxTile) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? (STAGE == 0 ? 5 : 7) : 1)))
void k_flux(ArenaPtrs<T> a) {
  constexpr bool BASE = STAGE == 0;
  const int2 blk = cell_block(a);
  const int tile = blk.x;
  if (tile < 0) return;
  const int m = blk.y;
  Control* c = a.ctrl + m;
  // The last stage pass commits the trial: every block of a slot with an active trial reports
  // (also when the trial has already failed), the last one commits.
  const bool commits = STAGE == 2 && a.commit_in_flux;
  if (commits ? !c->active : !pass_live(c, BASE ? PASS_BASE : PASS_STAGE)) return;
  const bool live = running(c);
  const MeshDesc<T>& d = a.desc[m];
  const int nc = a.nc;
  // tn cells per tile; nt = SPREAD * tn threads share the tile's face items, and thread tid < tn
  // owns cell first + tid.
  const int nt = static_cast<int>(blockDim.x), tid = static_cast<int>(threadIdx.x);
  const int tn = SPREAD == 1 ? nt : nt / SPREAD;
  const int first = tile * tn;
  const int k = first + tid;
  const bool valid = (SPREAD == 1 || tid < tn) && k < d.cell_begin + d.n_cells;
  const bool sa = d.sa != 0, laminar = d.laminar != 0;
  const int rb = read_buffer(c, STAGE);
  const int bit = BASE ? BAD_BASE : BAD_STAGE;
  const int staged = min(tn, d.cell_begin + d.n_cells - first);  // valid cells of the tile
  const TileCells<T> lds(tn);
  const FaceSlots<T> slots(tn, BASE);
  __shared__ int failed;  // block snapshot of the slot's admissibility bits (or the trial ended)
  if (tid == 0) failed = live ? c->bad : 1;
  // The cell's face-entry range is loaded with its record instead of after the barrier (stage
  // flux -2.5 to -3 us on 256 LF2 meshes).
  int csr_begin = 0, csr_end = 0;
  if (valid && live) {
    stage_cell(lds, tid, load_prim(a, rb, k));
    csr_begin = a.csr_off[k];
    csr_end = a.csr_off[k + 1];
  }
  __syncthreads();
  int bad = 0;  // admissibility bits this thread raises
  double local_dt = INFINITY;
  if (!failed) {
  // Face items of the tile, one chunk at a time: (A) each face touching the tile is evaluated
  // once, by one thread, into its LDS slot; (B) each cell sums its faces' slots in its cell-face
  // order.  A face shared with another tile is evaluated by both tiles from the same inputs, so
  // both cells see the same bits.
  const int2 range = a.item_range[tile];
  const int chunk = item_chunk(tn, BASE);
  T s0 = 0, s1 = 0, s2 = 0, s3 = 0, spectral = 0, diffusion = 0, sa_net = 0, sa_spec = 0;
  bool ok = true;
  for (int c0 = range.x; c0 < range.y; c0 += chunk) {
    const int c1 = min(range.y, c0 + chunk);
    // The base pass loads the next item's words while this one is evaluated (2 us less on 256
    // LF2 meshes); the stage passes, held to 72 registers, would spill them and lose 9-11 us.
    constexpr bool kPrefetch = BASE;
    ItemWords<T> next{};
    if (kPrefetch && c0 + tid < c1) next = load_item_words(a.ITEM_A, static_cast<size_t>(c0 + tid));
    for (int i = c0 + tid; i < c1; i += nt) {
      const ItemWords<T> words = kPrefetch ? next : load_item_words(a.ITEM_A, static_cast<size_t>(i));
      if (kPrefetch && i + nt < c1) next = load_item_words(a.ITEM_A, static_cast<size_t>(i + nt));
      const FaceItem<T> it = item_of(words, a.ITEM_B, static_cast<size_t>(i));
      CellPrim<T> L{}, R{};
      CellGrad<T> Lg{}, Rg{};
      const int ll = it.left - first, rl = it.right - first;
      if (it.left >= 0) {
        if (ll >= 0 && ll < staged) {
          L = tile_prim(lds, ll);
        } else {
          L = load_prim(a, rb, it.left);
        }
        if (laminar) Lg = load_grad(a, it.left);
      }
      if (it.right >= 0) {
        if (rl >= 0 && rl < staged) {
          R = tile_prim(lds, rl);
        } else {
          R = load_prim(a, rb, it.right);
        }
        if (laminar) Rg = load_grad(a, it.right);
      }
      // A boundary face has one cell: it is the inside cell and both sides' record.
      if (it.left < 0) {
        L = R;
        Lg = Rg;
      } else if (it.right < 0) {
        R = L;
        Rg = Lg;
      }
      const FaceFlux<T> f =
          face_flux<T, BASE>(d, make_int4(it.left, it.right, it.info, 0), it.geo, L, Lg, L, R, Lg, Rg);
      ok = ok && f.ok;
      const int sl = i - c0;
      store4(slots.f + 4 * sl, f.f0, f.fx, f.fy, f.f3);
      if (BASE) store4(slots.x + 4 * sl, f.speed * it.geo.len, f.fcoef * it.geo.len, f.sfcoef, f.sflux);
      else if (sa) slots.sf[sl] = f.sflux;
    }
    __syncthreads();
    if (valid) {
      // The cell's slot entries four at a time: the entries of a batch are loaded together
      // (entries past the end are loaded in range and skipped), then summed in cell-face order,
      // so the cell waits for one memory round trip per batch instead of one per face (stage flux
      // 157 / 167 -> 155 / 160 us on 256 LF2 meshes; base flux 151 -> 141 us once its register
      // bound no longer made it spill).
      const int lo = c0 - range.x, n_slots = c1 - c0;
      const int begin = csr_begin, end = csr_end;
      constexpr int kEntries = 4;
      for (int j0 = begin; j0 < end; j0 += kEntries) {
        int ent[kEntries];
#pragma unroll
        for (int q = 0; q < kEntries; ++q) ent[q] = a.SLOT[min(j0 + q, end - 1)];
#pragma unroll
        for (int q = 0; q < kEntries; ++q) {
        const int e = ent[q];
        const int sl = (e >> 1) - lo;
        if (j0 + q >= end || sl < 0 || sl >= n_slots) continue;
        const T sgn = (e & 1) ? T(-1) : T(1);
        T f0, fx, fy, f3;
        load4(slots.f + 4 * sl, f0, fx, fy, f3);
        s0 += sgn * f0;
        s1 += sgn * fx;
        s2 += sgn * fy;
        s3 += sgn * f3;
        if (BASE) {
          T xs, xd, xa, xf;
          load4(slots.x + 4 * sl, xs, xd, xa, xf);
          if (sa) sa_net += sgn * xf;
          spectral += xs;
          if (laminar) diffusion += xd;
          if (sa) sa_spec += xa;
        } else if (sa) {
          sa_net += sgn * slots.sf[sl];
        }
        }
      }
    }
    __syncthreads();  // the next chunk reuses the slots
  }
  if (!ok) bad |= bit;  // a face of the tile, whichever thread evaluated it
  ok = true;
  if (valid) {
    // The cell's own record is read back from LDS rather than kept in registers across the
    // face loop (fewer live registers there).
    const CellPrim<T> own = tile_prim(lds, tid);
    const T rho = own.rho, p = own.p;
    // The cell's own words, loaded together before the wall loop (each would otherwise wait for
    // the loads above it): wall range, inverse volume, SA point source; base pass: dt area,
    // volume, SA source rate.
    const int w0 = a.wall_first[k], w1 = w0 + a.wall_cnt[k];
    const T iv = a.inv_vol[k];
    const T sa_point_source = sa ? a.GREC[8 * k + 7] : T(0);
    const T dt_area = BASE ? a.dt_area[k] : T(0);
    const T cell_vol = BASE && (laminar || sa) ? a.vol[k] : T(0);
    const T src_rate = BASE && sa ? a.SRATE[k] : T(0);
    if (w1 > w0) {
      const T snd = sqrt(d.gamma * p / rho);
      for (int j = w0; j < w1; ++j) {
        const T len = a.wall_len[j];
        s1 += p * a.wall_nx[j] * len;
        s2 += p * a.wall_ny[j] * len;
        if (BASE) spectral += snd * len;
        if (laminar) {
          const int dn = a.wall_donor[j];
          const int dl = dn - first;
          const CellPrim<T> donor = dn == k ? own : (dl >= 0 && dl < staged ? tile_prim(lds, dl) : load_prim(a, rb, dn));
          const T dd = a.wall_ddist[j];
          T fv[4];
          wall_shear(donor, a.wall_nx[j], a.wall_ny[j], dd, fv);
          if (!isfinite(fv[1]) || !isfinite(fv[2])) ok = false;
          s1 -= fv[1] * len;
          s2 -= fv[2] * len;
          if (BASE) diffusion += T(4.0 / 3.0) * fmax(donor.mu / donor.rho, own.mu / rho) / dd * len;
          if (sa) {
            const T coeff = donor.mu * T(1.0 / SaConst<T>::sigma) / dd * len;
            sa_net -= -coeff * donor.phi;
            if (BASE) sa_spec += coeff;
          }
        }
      }
    }
    if (!ok) bad |= bit;
    const T r0 = -s0 * iv, r1 = -s1 * iv, r2 = -s2 * iv, r3 = -s3 * iv;
    T r4 = T(0);
    if (sa) {
      r4 = -sa_net * iv + sa_point_source;
      if (!isfinite(r4)) bad |= bit;
    }
    if (BASE) {
      a.R0[k] = r0;
      a.R0[nc + k] = r1;
      a.R0[2 * nc + k] = r2;
      a.R0[3 * nc + k] = r3;
      if (sa) a.R0[4 * nc + k] = r4;
      T dt = spectral > T(0) ? d.cfl * dt_area / spectral : T(INFINITY);
      if (laminar) {
        if (d.visc_mode == VISC_VOLUME) {
          if (diffusion > T(0)) dt = fmin(dt, d.transport.viscous_cfl * cell_vol / diffusion);
        } else if (d.visc_mode == VISC_SRD_RATE) {
          const T rate = diffusion * iv;
          if (rate > T(0)) dt = fmin(dt, d.transport.viscous_cfl / rate);
        } else {
          a.DRATE[k] = diffusion * iv;
        }
      }
      if (sa) {
        const T rate = T(2) * sa_spec / (rho * cell_vol) + src_rate;
        if (rate > T(0)) dt = fmin(dt, d.transport.sa_cfl / rate);
      }
      // SRD-grouped viscous limit: the cell's convective (and SA) limit waits for the group
      // rates (k_dt_group), as the reference's cell_dt does (src/static_amr2d_cuda.cu:159-161).
      if (laminar && d.visc_mode == VISC_SRD_GROUP) a.CELL_DT[k] = dt;
      local_dt = static_cast<double>(dt);
    } else if (stage_deltas(d)) {
      // SSP-RK3 on increments: U2 = U + D2 with D2 = (D1 + dt r) / 4, then
      // U' = U + 2/3 (D2 + dt r), the same combinations as below.  D1 = T(dt R0) is recomputed
      // here exactly as the first stage formed it (the base residual is intact), so the first
      // stage stores no increment.
      const double dt = c->trial_dt;
      const int cur = c->cur;
      const double* u = a.U[cur];
      const T rv[5] = {r0, r1, r2, r3, r4};
      double q[5] = {0, 0, 0, 0, 0};
      // Every load of the cell first (the fifth variable only with SA), then the sums and stores:
      // a loop over the run-time variable count that stores as it goes waits for each load.
      T pv[5] = {0, 0, 0, 0, 0};
      double uv[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        if (v == 4 && d.nv != 5) break;
        const int i = v * nc + k;
        pv[v] = STAGE == 1 ? ld_last(a.R0 + i) : ld_last(a.DU + i);
        uv[v] = STAGE == 1 ? u[i] : ld_last(u + i);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        if (v == 4 && d.nv != 5) break;
        const int i = v * nc + k;
        const T prev = STAGE == 1 ? static_cast<T>(dt * static_cast<double>(pv[v])) : pv[v];
        const double inc = static_cast<double>(prev) + dt * static_cast<double>(rv[v]);
        if (STAGE == 1) {
          const T dn = static_cast<T>(0.25 * inc);
          a.DU[i] = dn;
          q[v] = uv[v] + static_cast<double>(dn);
        } else {
          q[v] = uv[v] + (2.0 / 3.0) * inc;
          a.U[cur ^ 1][i] = q[v];
        }
      }
      CellPrim<T> w;
      if (!make_prim_q(q, d, w)) bad |= BAD_STAGE;
      else store_prim(a, write_buffer(c, STAGE), k, w);
    } else {
      // SSP-RK3: u2 = 3/4 u + 1/4 (u1 + dt r)  or  u1 = 1/3 u + 2/3 (u2 + dt r).
      const double dt = c->trial_dt;
      const int cur = c->cur;
      const double* u = a.U[cur];
      double* out = STAGE == 1 ? a.U2 : a.U[cur ^ 1];
      const double* prev = STAGE == 1 ? a.U[cur ^ 1] : a.U2;
      const double wu = STAGE == 1 ? 0.75 : 1.0 / 3.0, wp = STAGE == 1 ? 0.25 : 2.0 / 3.0;
      // Every load first, then the combinations and stores, then the primitives from registers
      // (a store before the next variable's loads makes each load wait for the one before).
      double uv[5] = {0, 0, 0, 0, 0}, pv[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        uv[v] = u[v * nc + k];
        pv[v] = prev[v * nc + k];
      }
      if (sa) {
        uv[4] = u[4 * nc + k];
        pv[4] = prev[4 * nc + k];
      }
      const T rv[5] = {r0, r1, r2, r3, r4};
      double q[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        if (v == 4 && !sa) break;
        q[v] = wu * uv[v] + wp * (pv[v] + dt * static_cast<double>(rv[v]));
        out[v * nc + k] = q[v];
      }
      if (!d.small) {
        CellPrim<T> w;
        if (!make_prim_q(q, d, w)) bad |= BAD_STAGE;
        else store_prim(a, write_buffer(c, STAGE), k, w);
      }
    }
  }
  }  // !failed
  if (BASE && d.visc_mode != VISC_SRD_GROUP) publish_dt(c, local_dt);
  if (commits) {
    const int any = __syncthreads_or(bad);
    if (tid == 0) arrive_and_commit(c, d.n_tiles, any != 0);
  } else if (bad) {
    flag_bad(c, bad);
  }
}

// SRD-grouped viscous limit: group means of the diffusion rate, then the per-cell limit
// (reference diffusion_group_kernel and choose_dt SRD branch, src/static_amr2d_cuda.cu:164-176).
template <class T>
__global__ __launch_bounds__(kMaxTile) void k_diffusion_group(ArenaPtrs<T> a) {
  const int2 blk = cell_block(a);
  const int tile = blk.x;
  if (tile < 0) return;
  const int m = blk.y;
  Control* c = a.ctrl + m;
  if (!pass_live(c, PASS_BASE)) return;  // uniform over the block: it meets a barrier in publish_dt
  const MeshDesc<T>& d = a.desc[m];
  if (d.visc_mode != VISC_SRD_GROUP) return;
  const int g = tile * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
  if (g >= d.cell_begin + d.n_cells) return;
  T sum = 0;
  for (int j = a.grp_off[g]; j < a.grp_off[g + 1]; ++j) sum += a.grp_w[j] * a.DRATE[a.grp_mem[j]];
  a.DGROUP[g] = sum;
}

template <class T>
__global__ __launch_bounds__(kMaxTile) void k_dt_group(ArenaPtrs<T> a) {
  const int2 blk = cell_block(a);
  const int tile = blk.x;
  if (tile < 0) return;
  const int m = blk.y;
  Control* c = a.ctrl + m;
  if (!pass_live(c, PASS_BASE)) return;  // uniform over the block: it meets a barrier in publish_dt
  const MeshDesc<T>& d = a.desc[m];
  if (d.visc_mode != VISC_SRD_GROUP) return;
  const int k = tile * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
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
// Step size of the trial and the first RK stage (reference choose_dt_kernel and rk_kernel,
// src/static_amr2d_cuda.cu:177-194).  (Opening the first stage inside the stage-1 gradient pass
// instead, which then forms the first-stage velocities of neighbours outside its tile from U and
// R0 itself, was measured and dropped: 146.8 us against 49.4 + 79.8 us for the two launches on
// 256 LF2 meshes, and no gain in the latency of a single-mesh trial, 70.5 against 70.4 us.)
template <class T>
__global__ __launch_bounds__(kMaxTile) void k_stage0(ArenaPtrs<T> a) {
  const int2 blk = cell_block(a);
  const int tile = blk.x;
  if (tile < 0) return;
  const int m = blk.y;
  Control* c = a.ctrl + m;
  if (!running(c)) return;
  const MeshDesc<T>& d = a.desc[m];
  double dt;
  if (!open_stage0(c, d, tile, &dt)) return;
  const int k = tile * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
  if (k >= d.cell_begin + d.n_cells) return;
  // U1 = U + D1, D1 = T(dt R0) not stored (the next stage re-forms it); or U1 in U[cur ^ 1].
  double q[5];
  first_stage(a, d, c, k, dt, true, q);
  if (d.small && !stage_deltas(d)) return;  // primitives after the redistribution
  CellPrim<T> w;
  if (!make_prim_q(q, d, w)) flag_bad(c, BAD_STAGE);
  else store_prim(a, write_buffer(c, 0), k, w);
}

// SRD: q_g = sum_k w_gk u_k, then u_k = mean of q over the groups holding k, then primitives.
// `stage` selects the candidate: 0 and 2 write U[cur ^ 1], 1 writes U2.
template <class T>
__global__ __launch_bounds__(kMaxTile) void k_group(ArenaPtrs<T> a, int stage) {
  const int2 blk = cell_block(a);
  const int tile = blk.x;
  if (tile < 0) return;
  const int m = blk.y;
  Control* c = a.ctrl + m;
  if (!healthy(c)) return;
  const MeshDesc<T>& d = a.desc[m];
  if (!d.small) return;
  const int g = tile * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
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
  const int2 blk = cell_block(a);
  const int tile = blk.x;
  if (tile < 0) return;
  const int m = blk.y;
  Control* c = a.ctrl + m;
  if (!healthy(c)) return;
  const MeshDesc<T>& d = a.desc[m];
  if (!d.small) return;
  const int k = tile * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
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
  CellPrim<T> w;
  if (!make_prim(u, nc, k, d, w)) flag_bad(c, BAD_STAGE);
  else store_prim(a, write_buffer(c, stage), k, w);
}

// ------------------------------------------------------------------------------------------
// Commits slots first, first + stride, ... (< n): a branch of the trial commits its own slots.
__global__ void k_commit(Control* ctrl, int n, int first, int stride) {
  const int m = first + stride * static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (m >= n) return;
  Control* c = ctrl + m;
  commit_trial(c, c->bad != 0);
}

__global__ void k_target(Control* ctrl, int n, double target, long long maxsteps, const double* targets,
                         const long long* maxsteps_each) {
  const int m = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
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
  const int2 blk = cell_block(a);
  const int tile = blk.x;
  if (tile < 0) return;
  const int m = blk.y;
  const MeshDesc<T>& d = a.desc[m];
  const int k = tile * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
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
    for (int i = 0; i < 6; ++i) a.MON_CELL[6 * tile + i] = acc[i];
}

// Steady monitor, walls: pressure force minus viscous shear and the moment about (mx, my)
// (reference src/quick_airfoil.cpp:83-85), from the records of the accepted state.
template <class T>
__global__ __launch_bounds__(kMaxTile) void k_monitor_walls(ArenaPtrs<T> a) {
  const int b = blockIdx.x;
  const int m = a.face_block_mesh[a.n_face_blocks + b];
  const MeshDesc<T>& d = a.desc[m];
  const int j = b * blockDim.x + threadIdx.x;
  double acc[3] = {0, 0, 0};
  if (j < d.wall_begin + d.n_walls) {
    const int wb = a.ctrl[m].wbuf;
    const double mx = a.mon_frame[6 * m + 4], my = a.mon_frame[6 * m + 5];
    const CellPrim<T> own = load_prim(a, wb, a.wall_owner[j]);
    const T len_t = a.wall_len[j];
    const double p = static_cast<double>(own.p), len = static_cast<double>(len_t);
    double fx = p * static_cast<double>(a.wall_nx[j]) * len, fy = p * static_cast<double>(a.wall_ny[j]) * len;
    if (d.laminar) {
      const CellPrim<T> donor = load_prim(a, wb, a.wall_donor[j]);
      T fv[4];
      wall_shear(donor, a.wall_nx[j], a.wall_ny[j], a.wall_ddist[j], fv);
      fx -= static_cast<double>(fv[1] * len_t);
      fy -= static_cast<double>(fv[2] * len_t);
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

#define HIPAMR_LAUNCH_LDS(name, kernel, grid, lds, ...)             \
  do {                                                               \
    if (hook) hook->before(name);                                    \
    hipLaunchKernelGGL(kernel, grid, tile, lds, s, __VA_ARGS__);     \
    if (hook) hook->after(name);                                     \
    ++launches;                                                      \
  } while (0)
#define HIPAMR_LAUNCH(name, kernel, grid, ...) HIPAMR_LAUNCH_LDS(name, kernel, grid, 0, __VA_ARGS__)
// A flux launch of stage STAGE: the tile's LDS, flux_spread threads per cell.
#define HIPAMR_LAUNCH_FLUX(name, STAGE, grid, lds, ...)                                                   \
  do {                                                                                                   \
    if (hook) hook->before(name);                                                                        \
    if (a.flux_spread == 2)                                                                              \
      hipLaunchKernelGGL((k_flux<T, STAGE, 2>), grid, dim3(2 * a.tile), lds, s, __VA_ARGS__);          \
    else                                                                                                 \
      hipLaunchKernelGGL((k_flux<T, STAGE, 1>), grid, dim3(a.tile), lds, s, __VA_ARGS__);              \
    if (hook) hook->after(name);                                                                         \
    ++launches;                                                                                          \
  } while (0)

// The base pass proper (gradients, faces, residual, local dt); the base primitives are valid.
template <class T>
int launch_base_pass(const ArenaPtrs<T>& a, hipStream_t s, LaunchHook* hook) {
  const dim3 tile(a.tile), cells(a.n_grid_blocks);
  const size_t base_lds = flux_lds_bytes<T>(a.tile, true), vel_lds = grad_lds_bytes<T>(a.tile);
  int launches = 0;
  if (a.any_laminar) HIPAMR_LAUNCH_LDS("gradient_base", (k_grad<T, 0>), cells, vel_lds, a);
  HIPAMR_LAUNCH_FLUX("flux_base", 0, cells, base_lds, a);
  if (a.any_visc_group) {
    HIPAMR_LAUNCH("diffusion_group", k_diffusion_group<T>, cells, a);
    HIPAMR_LAUNCH("dt_group", k_dt_group<T>, cells, a);
  }
  return launches;
}

template <class T>
int launch_prims(const ArenaPtrs<T>& a, hipStream_t s, LaunchHook* hook) {
  const dim3 tile(a.tile), cells(a.n_grid_blocks);
  int launches = 0;
  HIPAMR_LAUNCH("base_primitive", k_base_prim<T>, cells, a);
  return launches;
}

template <class T>
int launch_residual(const ArenaPtrs<T>& a, hipStream_t s, LaunchHook* hook) {
  return launch_prims(a, s, hook) + launch_base_pass(a, s, hook);
}

// A trial never needs the base primitives recomputed: they are valid after launch_prims (run
// whenever the host starts trials) and survive rejected trials.
template <class T>
int launch_trial(const ArenaPtrs<T>& a, hipStream_t s, LaunchHook* hook) {
  const dim3 tile(a.tile), cells(a.n_grid_blocks);
  const size_t stage_lds = flux_lds_bytes<T>(a.tile, false), vel_lds = grad_lds_bytes<T>(a.tile);
  int launches = launch_base_pass(a, s, hook);
  HIPAMR_LAUNCH("stage0", k_stage0<T>, cells, a);
  if (a.any_small) {
    HIPAMR_LAUNCH("srd_group", k_group<T>, cells, a, 0);
    HIPAMR_LAUNCH("srd_redistribute", k_redist<T>, cells, a, 0);
  }
  if (a.any_laminar) HIPAMR_LAUNCH_LDS("gradient_stage", (k_grad<T, 1>), cells, vel_lds, a);
  HIPAMR_LAUNCH_FLUX("flux_rk1", 1, cells, stage_lds, a);
  if (a.any_small) {
    HIPAMR_LAUNCH("srd_group", k_group<T>, cells, a, 1);
    HIPAMR_LAUNCH("srd_redistribute", k_redist<T>, cells, a, 1);
  }
  if (a.any_laminar) HIPAMR_LAUNCH_LDS("gradient_stage", (k_grad<T, 2>), cells, vel_lds, a);
  HIPAMR_LAUNCH_FLUX("flux_rk2", 2, cells, stage_lds, a);
  if (a.any_small) {
    HIPAMR_LAUNCH("srd_group", k_group<T>, cells, a, 2);
    HIPAMR_LAUNCH("srd_redistribute", k_redist<T>, cells, a, 2);
  }
  if (!a.commit_in_flux) {
    const dim3 tile(256);
    const int count = (a.n_meshes - a.commit_first + a.commit_stride - 1) / a.commit_stride;
    HIPAMR_LAUNCH("commit", k_commit, dim3(blocks_of(count, 256)), a.ctrl, a.n_meshes, a.commit_first,
                  a.commit_stride);
  }
  return launches;
}
#undef HIPAMR_LAUNCH
#undef HIPAMR_LAUNCH_LDS
#undef HIPAMR_LAUNCH_FLUX

template <class T>
void launch_monitor(const ArenaPtrs<T>& a, hipStream_t s) {
  hipLaunchKernelGGL(k_monitor_cells<T>, dim3(a.n_grid_blocks), dim3(a.tile), 0, s, a);
  if (a.n_wall_blocks > 0) hipLaunchKernelGGL(k_monitor_walls<T>, dim3(a.n_wall_blocks), dim3(a.tile), 0, s, a);
}

}  // namespace hipamr

#endif
