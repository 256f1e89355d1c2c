// amr_arena.h -- device layout of one or many meshes advanced in lockstep.
//
// Every mesh of an arena owns block-aligned ranges of a global cell, face and wall index
// space (each range padded to the launch tile, so a thread block never spans two meshes).
// All index arrays hold global ids.  Per-element data are structure-of-arrays planes (cell
// planes have stride nc, face planes nf, wall planes nw), except what kernels gather from
// neighbours, which is kept as short per-element vectors (VREC, PREC, GREC, face items).
//
// Per mesh, an on-device Control drives the SSP-RK3 trial state machine of the reference
// (src/static_amr2d_cuda.cu:17-21,51-59,170-187,212-216): time, target, dt, step and
// rejection counters, the admissibility flag and fatal codes 2 (nonphysical base / bad at dt
// choice), 3 (maxsteps), 4 (dt underflow), 5 (21 rejections).  The accepted state alternates
// between two buffers (Control::cur), so no commit copy is needed.
#ifndef HIPAMR_AMR_ARENA_H
#define HIPAMR_AMR_ARENA_H

#include <cstdint>

#include "amr_physics.h"

namespace hipamr {

struct Control {
  double time, target, dt;   // dt: step size of the last trial (halved on rejection)
  double trial_dt;           // step size of the running trial
  unsigned long long dt_min; // bit pattern of the minimum local dt of the base pass
  long long steps, rejected, maxsteps;
  int bad, fatal, retrying, attempt, finished, active;
  int cur;                   // buffer holding the accepted state: U[cur]
  int w_valid;               // W/PHI hold primitives of U[cur]
  int r0_valid;              // R0 holds L(U[cur]) (set after a rejection)
  int wbuf;                  // primitive-vector buffer (0..2) holding the primitives of U[cur]
  // Blocks of the last stage pass that have finished this trial (low 32 bits) and how many of
  // them found the candidate inadmissible (high 32 bits); the last one commits (zero between
  // trials).
  unsigned long long arrivals;
};
static_assert(sizeof(Control) == 112, "Control layout");

enum : int { VISC_VOLUME = 0, VISC_SRD_RATE = 1, VISC_SRD_GROUP = 2 };

// Words per cell of the PREC vectors (p, mu, eddy viscosity): unpadded in FP32 (12 bytes, one
// dwordx3 access), padded to two 16-byte halves in FP64.
template <class T>
__host__ __device__ constexpr int prec_width() { return sizeof(T) == 4 ? 3 : 4; }

// Face items pack (right cell + 1) into the low bits of one int, the face info (4 bits) above
// them and the irregular-geometry flag in bit 31.
constexpr int kItemRightBits = 27;
constexpr int kItemInfoMask = 15;

// Face items held in LDS at once per block: 5/2 per cell of the tile (a tile of an LF1 / LF2
// mesh touches 2.2 - 2.7 faces per cell; a tile with more items is processed in several chunks).
// Measured on 256 LF2 meshes at tile 128 with 8-word base items: base chunks of 3/2, 2, 9/4 - 1/32,
// 5/2, 11/4 and 3 per cell give 164, 164, 178, 159, 171 and 171 us for the base pass and 700,
// 699, 705, 683, 690 and 691 us per batch step (one chunk round for most tiles against losing a
// block per CU); a stage chunk of 11/4 gains nothing, 3 loses 13 us per stage pass.
__host__ __device__ constexpr int item_chunk(int tile, bool) { return tile * 5 / 2; }

// One LSQ gradient term of a cell: its source (>= 0 cell, -1 body wall, -2 - side: no-slip domain
// side) and its two weights, fetched by one load (12 bytes in FP32) instead of three.
template <class T>
struct TermRec;
template <>
struct TermRec<float> {
  int src;
  float wx, wy;
};
template <>
struct TermRec<double> {
  double wx, wy;
  int src, pad;
};

// Dynamic LDS of a flux launch: 8 T per cell (primitive vectors), then per item of a chunk 4 + 1
// T (convective and viscous flux, SA flux) in a stage pass, 4 + 4 T (fluxes; dt terms and SA
// flux) in the base pass.
// Gradient launches: 4 T per cell.
template <class T>
constexpr size_t flux_lds_bytes(int tile, bool base) {
  return (static_cast<size_t>(tile) * 8 + static_cast<size_t>(item_chunk(tile, base)) * (base ? 8 : 5)) * sizeof(T);
}
template <class T>
constexpr size_t grad_lds_bytes(int tile) {
  return static_cast<size_t>(tile) * 4 * sizeof(T);
}

template <class T>
struct MeshDesc {
  int cell_begin, n_cells, face_begin, n_faces, wall_begin, n_walls;
  int first_cell_block;  // block index (in cell launches) of the mesh's first tile
  int n_tiles;           // tiles of the slot: blocks of a cell launch that serve it
  int nv;                // 4, or 5 with SA-neg
  int laminar, sa, flux, small, visc_mode, agglomerated;
  double dt_fixed;
  T gamma, cfl, mco;
  Transport<T> transport;
  Side<T> side[4];
};

// Device pointers of an arena (passed by value to every kernel).
template <class T>
struct ArenaPtrs {
  int n_meshes, nc, nf, nw, tile;
  // Threads per cell of the flux launches (1 or 2; tile * flux_spread <= 256): the face items of
  // a tile are spread over flux_spread * tile threads, the cell work stays on the first tile.
  int flux_spread;
  int n_cell_blocks, n_face_blocks, n_wall_blocks;
  int n_grid_blocks;  // blocks of a cell launch (>= n_cell_blocks, a multiple of 8)
  int commit_first, commit_stride;  // slots a trial's commit covers: first, first + stride, ...
  int any_laminar, any_sa, any_small, any_visc_group;
  // The last stage pass commits the trial (its last block per slot); without SRD meshes only,
  // whose last stage is redistributed after it (then k_commit follows the redistribution).
  int commit_in_flux;
  const MeshDesc<T>* desc;
  Control* ctrl;
  const int2* block_rec;       // n_grid_blocks: {tile, slot} of each block of a cell launch, {-1, -1} none
  const int* face_block_mesh;  // n_face_blocks + n_wall_blocks (faces first, then walls)
  // cells
  const T* vol;
  const T* inv_vol;
  const T* dt_area;
  const T* wall_dist;
  const int* csr_off;    // nc + 1
  const uint16_t* SLOT;  // per cell-face entry: index of the face among its tile's items << 1,
                         // | 1 when the cell is the face's right cell
  const int* wall_first; // first wall of the cell
  const int* wall_cnt;
  const int* term_off;   // nc + 1
  const TermRec<T>* terms;  // LSQ gradient terms in the cells' term order
  // SRD neighbourhoods: group g = cell g
  const int* grp_off;    // nc + 1
  const int* grp_mem;
  const T* grp_w;        // V_k / (N_k Vhat_g)
  const int* mem_off;    // nc + 1
  const int* mem_grp;
  const T* mem_inv;      // 1 / N_k
  // Face items: per cell block (tile) the faces with at least one cell in the tile, [x, y) in
  // the item arrays, 4 T words per item in each.  ITEM_A: ints {left or -1, (right + 1) |
  // info << kItemRightBits | irregular << 31} (words 0-1 in FP32, word 0 in FP64), length, d_n;
  // ITEM_B: d_t, ex, ey, distance, where d_n / d_t are the centroid offsets along / across the
  // face's axis.  Face info: bit 0 axis, bits 1..3 side + 1 (0 interior).  A regular item
  // (d_t = ex = ey = 0 and distance = |d_n|, bit for bit) has its ITEM_B words implied and never
  // read; a tile lists its irregular items first, so the ITEM_B words it reads lie together.
  const int2* item_range;
  const T* ITEM_A;
  const T* ITEM_B;
  // walls
  const int* wall_owner;
  const int* wall_donor;
  const T* wall_nx;
  const T* wall_ny;
  const T* wall_len;
  const T* wall_ddist;
  // steady monitor: FP64 centroids and fragment midpoints, per-block FP64 partial sums
  const double* cell_cx;
  const double* cell_cy;
  const double* wall_mx;
  const double* wall_my;
  const double* mon_frame;  // n_meshes x 6: region xmin, xmax, ymin, ymax; moment point x, y
  double* MON_CELL;  // n_cell_blocks x 6: sum V|r_v| (v < 5) and sum V over the region
  double* MON_WALL;  // n_wall_blocks x 3: fx, fy, mz
  // Conserved state: FP64 in both precisions (the RK update dt*r must stay above one ulp of u
  // down to the steady tolerance); residuals and all work arrays are T.
  double* U[2];
  double* U2;
  T* DU;         // nv planes: RK stage increments over U[cur] (FP32 work precision without SRD)
  double* QHAT;  // SRD group averages, 4 planes
  T* R0;
  // Cell vectors gathered by faces and gradient stencils (AoS per purpose, so that a gather
  // costs whole vector loads): VREC = rho, u, v, nutilde and PREC = p, mu, eddy viscosity
  // (prec_width T per cell: 3 in FP32, a 12-byte load; 4 in FP64, two 16-byte loads) (three
  // buffers rotated by stage, Control::wbuf holding U[cur]'s), GREC = ux, uy, vx, vy,
  // nutilde_x, nutilde_y, SA diffusion, SA source (per stage).
  T* VREC[3];
  T* PREC[3];
  T* GREC;
  T* SRATE;   // SA-neg source rate (base pass; the source itself rides in GREC's 8th slot)
  T* CELL_DT; // deferred local dt (SRD-grouped viscous limit only)
  T* DRATE;
  T* DGROUP;
};

// Optional observer of the launch sequence (per-kernel timing): before(name) is called right
// before and after(name) right after each kernel launch on the stream.
struct LaunchHook {
  virtual void before(const char* name) = 0;
  virtual void after(const char* name) = 0;
  virtual ~LaunchHook() = default;
};

// Launch one full trial (one SSP-RK3 step attempt) of every mesh on `stream`.
// Returns the number of kernel launches it issued.
template <class T>
int launch_trial(const ArenaPtrs<T>& a, hipStream_t stream, LaunchHook* hook = nullptr);
// Base pass only: primitives of U[cur] and R0 = L(U[cur]) (the residual evaluation).
template <class T>
int launch_residual(const ArenaPtrs<T>& a, hipStream_t stream, LaunchHook* hook = nullptr);
// Primitives of U[cur] for every running mesh whose base primitives are not valid (after the
// host wrote its state): must follow every launch_target before trials are launched.
template <class T>
int launch_prims(const ArenaPtrs<T>& a, hipStream_t stream, LaunchHook* hook = nullptr);
// Steady-monitor partial sums of the residual R0 (cells whose centroid lies in the mesh's
// region) and of the wall pressure/shear force and moment about the mesh's moment point, both
// taken from mon_frame.
template <class T>
void launch_monitor(const ArenaPtrs<T>& a, hipStream_t stream);
// Set the target time and step cap of every mesh and start a trial.  targets/maxsteps are
// device arrays of n_meshes entries, or null to use the scalar values for all meshes.
void launch_target(Control* ctrl, int n_meshes, double target, long long maxsteps, const double* targets,
                   const long long* maxsteps_each, hipStream_t stream);

}  // namespace hipamr

#endif
