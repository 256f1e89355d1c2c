// amr_device.hip -- slot layout of an arena, one-allocation upload, admission, state transfer.
//
// Every array lives in one slab (256-byte aligned sub-allocations), hipMalloc'ed by the arena or
// handed to it (make_arena's `slab`).  Host copies
// of the geometry tables are kept so that admitting a mesh into a slot rewrites and uploads
// only that slot's ranges.  Index spaces and their per-slot ranges:
//   cells   [cell_begin, cell_begin + cap.cells)   (tile multiple, >= n_cells + 1)
//   faces   [face_begin, face_begin + cap.faces)   walls [wall_begin, wall_begin + cap.walls)
//   csr     [csr_begin, csr_begin + cap.csr)       terms [term_begin, term_begin + cap.terms)
//   groups  [grp_begin, grp_begin + cap.groups)    members [mem_begin, mem_begin + cap.members)
// Offsets arrays (nc + 1 entries) hold global positions; a slot's padding cells repeat the
// end of its used range, so a real cell's range never reaches unused entries.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "amr_device.h"

namespace hipamr {
namespace {

constexpr int BC_TRANSMISSIVE = 1, BC_REFLECT = 2, BC_FARFIELD = 3, BC_TOTAL_INLET = 4, BC_PRESSURE_OUTLET = 5,
              BC_NOSLIP = 6;
constexpr unsigned long long kInf = 0x7ff0000000000000ull;

// Zero `n` 32-bit words (grid-stride): the one-time clear of an arena's work arrays.
__global__ void k_clear_words(uint32_t* p, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

bool state_ok(const double* w) {
  return std::isfinite(w[0]) && w[0] > 0 && std::isfinite(w[1]) && std::isfinite(w[2]) && std::isfinite(w[3]) &&
         w[3] > 0;
}

// Reference amr_bc_valid (src/static_amr2d_boundary.h:32-47).
bool side_valid(const amr_boundary& b, int legacy, int side) {
  if (b.kind < -1) return false;
  const int kind = b.kind < 0 ? legacy : b.kind;
  if (kind == BC_TRANSMISSIVE || kind == BC_REFLECT) return true;
  if (kind == BC_NOSLIP)
    return std::isfinite(b.wall_velocity[0]) && std::isfinite(b.wall_velocity[1]) && b.wall_velocity[side / 2] == 0;
  if (kind < BC_FARFIELD || kind > BC_PRESSURE_OUTLET || !state_ok(b.reference)) return false;
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

// Reference amr_transport_valid (src/static_amr2d_viscous.h:148-156).
bool transport_ok(const amr_transport& t) {
  if (t.model == 0) return t.sa_model == 0;
  return (t.sa_model == 0 || (t.sa_model == 1 && std::isfinite(t.sa_nu_inf) && t.sa_nu_inf >= 0 &&
                              std::isfinite(t.sa_cfl) && t.sa_cfl > 0 && t.sa_cfl <= .1)) &&
         t.model == 1 && (t.law == 0 || t.law == 1) && std::isfinite(t.mu_ref) && t.mu_ref > 0 &&
         std::isfinite(t.gas_constant) && t.gas_constant > 0 && std::isfinite(t.temperature_ref) &&
         t.temperature_ref > 0 && std::isfinite(t.sutherland) && t.sutherland >= 0 && std::isfinite(t.viscous_cfl) &&
         t.viscous_cfl > 0 && t.viscous_cfl <= .2 && (t.stabilization == 0 || t.stabilization == 1) &&
         std::isfinite(t.wall_probe_min_h) && t.wall_probe_min_h >= 0 && t.wall_probe_min_h <= 2;
}

// Reference static_amr2d_viscous_ready (src/static_amr2d_viscous.cpp:34-38) on a view.
bool tables_ready(const amr_mesh_view& v) {
  const amr_transport& t = v.transport;
  if (t.model == 0)
    for (const amr_boundary& b : v.boundaries)
      if (b.kind == BC_NOSLIP) return false;
  if (!transport_ok(t)) return false;
  if (t.sa_model && !v.wall_distance) return false;
  if (t.model == 1 && (!v.visc_offsets || (v.n_walls && !v.wall_samples) || !v.face_samples || !v.face_geometry))
    return false;
  return true;
}

inline int round_up(int n, int tile) { return (n + tile - 1) / tile * tile; }
inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Element counts of an arena's index spaces.
struct Totals {
  size_t meshes, cells, faces, walls, csr, terms, groups, members, cell_blocks, face_blocks, wall_blocks;
  size_t grid_blocks;
};

// Tile of each block of a cell launch.  Blocks go to the 8 XCDs in turn (block b to XCD b % 8);
// every slot's tiles are cut into 8 contiguous runs, run j going to XCD (j + slot) % 8, and each
// XCD takes its runs in slot order.  So any set of running slots keeps all XCDs busy, tiles that
// share neighbours share an L2, and the rotation evens out the runs' unequal lengths.  Blocks
// past the end of an XCD's list get -1.
// With branches > 1 only the slots m with m % branches == branch are listed (the table of one
// branch of the trial, see branch_count).
std::vector<int> block_tiles(const std::vector<SlotCapacity>& caps, int tile, int branches = 1, int branch = 0) {
  std::vector<int> list[8];
  int t0 = 0;
  for (size_t m = 0; m < caps.size(); ++m) {
    const int n = caps[m].cells / tile;
    if (static_cast<int>(m % branches) != branch) {
      t0 += n;
      continue;
    }
    for (int j = 0; j < 8; ++j) {
      const int lo = t0 + static_cast<int>(static_cast<long long>(j) * n / 8);
      const int hi = t0 + static_cast<int>(static_cast<long long>(j + 1) * n / 8);
      std::vector<int>& l = list[(j + m) % 8];
      for (int t = lo; t < hi; ++t) l.push_back(t);
    }
    t0 += n;
  }
  size_t longest = 0;
  for (const auto& l : list) longest = std::max(longest, l.size());
  std::vector<int> out(8 * std::max<size_t>(longest, 1), -1);
  for (int x = 0; x < 8; ++x)
    for (size_t i = 0; i < list[x].size(); ++i) out[8 * i + x] = list[x][i];
  return out;
}

// An arena of several slots runs its trial as independent branches (slot m in branch
// m % branches) on their own streams, so that one branch's kernels fill the GPU while another's
// drain.  Three measured best for 256 LF2 meshes (703.1 us per step, against 715.7 with two and
// 710.8 with four; 951 with one stream in an earlier build).
int branch_count(size_t slots) { return static_cast<int>(std::min<size_t>(slots, 3)); }

// The block tables of an arena: the whole arena first, then one per branch (when branched).
std::vector<int> all_block_tiles(const std::vector<SlotCapacity>& caps, int tile, std::vector<int>* offsets) {
  std::vector<int> out = block_tiles(caps, tile);
  offsets->assign(1, 0);
  const int k = branch_count(caps.size());
  if (k > 1)
    for (int j = 0; j < k; ++j) {
      offsets->push_back(static_cast<int>(out.size()));
      const std::vector<int> part = block_tiles(caps, tile, k, j);
      out.insert(out.end(), part.begin(), part.end());
    }
  offsets->push_back(static_cast<int>(out.size()));
  return out;
}

Totals totals_of(const std::vector<SlotCapacity>& caps, int tile) {
  Totals t{};
  t.meshes = caps.size();
  for (const SlotCapacity& c : caps) {
    t.cells += c.cells;
    t.faces += c.faces;
    t.walls += c.walls;
    t.csr += c.csr;
    t.terms += c.terms;
    t.groups += c.groups;
    t.members += c.members;
  }
  t.cell_blocks = t.cells / tile;
  std::vector<int> offsets;
  t.grid_blocks = all_block_tiles(caps, tile, &offsets).size();
  t.face_blocks = t.faces / tile;
  t.wall_blocks = t.walls / tile;
  return t;
}

// Byte plan of the slab: geometry tables, control records, work planes.  Every array takes at
// least one element and starts on a 256-byte boundary.
struct SlabPlan {
  size_t geometry, control, work, total;
  size_t cplane, dplane, vec4, prec, vec8, mon_cell, mon_wall;
};

template <class T>
SlabPlan plan_slab(const Totals& t, int nv) {
  const auto arr = [](size_t elem, size_t count) { return align256(elem * std::max<size_t>(count, 1)); };
  SlabPlan p{};
  p.geometry = arr(sizeof(MeshDesc<T>), t.meshes) + arr(8, t.grid_blocks) +
               arr(4, t.face_blocks + t.wall_blocks) +
               5 * arr(sizeof(T), t.cells) + 2 * arr(4, t.cells) + 2 * arr(8, t.cells) + 4 * arr(4, t.cells + 1) +
               arr(2, t.csr) + 2 * arr(4 * sizeof(T), t.csr) + arr(sizeof(TermRec<T>), t.terms) +
               arr(4, t.groups) + arr(sizeof(T), t.groups) + arr(4, t.members) + arr(8, t.cell_blocks) +
               2 * arr(4, t.walls) + 4 * arr(sizeof(T), t.walls) + 2 * arr(8, t.walls);
  p.control = arr(sizeof(Control), t.meshes) + arr(sizeof(double), t.meshes) + arr(sizeof(long long), t.meshes) +
              arr(6 * sizeof(double), t.meshes);
  p.cplane = arr(sizeof(T), t.cells);
  p.dplane = arr(sizeof(double), t.cells);
  p.vec4 = arr(4 * sizeof(T), t.cells);
  p.prec = arr(prec_width<T>() * sizeof(T), t.cells);
  p.vec8 = arr(8 * sizeof(T), t.cells);
  p.mon_cell = arr(6 * sizeof(double), t.cell_blocks);
  p.mon_wall = arr(3 * sizeof(double), t.wall_blocks);
  // FP64 state U[2], U2 (nv planes each) and QHAT (4); cell vectors VREC[3] (4 T), PREC[3]
  // (3 T in FP32, 4 in FP64) and GREC (8 T); T planes R0 (nv), SRATE, CELL_DT, DRATE, DGROUP, and with FP32 work
  // precision the RK stage increments DU (nv); monitor partials.
  const size_t state_planes = 3 * nv + 4, cell_planes = nv + 4 + (sizeof(T) == 4 ? nv : 0);
  p.work = state_planes * p.dplane + 3 * p.vec4 + 3 * p.prec + p.vec8 + cell_planes * p.cplane + p.mon_cell + p.mon_wall;
  p.total = p.geometry + p.control + p.work;
  return p;
}

// Distance of (x, y) along the Hilbert curve of a 2^16 x 2^16 grid (the classical iterative
// xy -> d conversion: per bit, pick the quadrant, then rotate / reflect the frame).
uint64_t hilbert_index(uint32_t x, uint32_t y) {
  uint64_t d = 0;
  for (uint32_t s = 1u << 15; s > 0; s >>= 1) {
    const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += static_cast<uint64_t>(s) * s * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = 0xFFFFu - x;
        y = 0xFFFFu - y;
      }
      const uint32_t t = x;
      x = y;
      y = t;
    }
  }
  return d;
}

// Storage order of a mesh on the device: cells along the Hilbert curve of their centroids
// (neighbours land in nearby tiles and cache lines), faces by the smaller storage index of their
// cells.  cell_pos[old] = storage position, cell_at[pos] = old; face_pos / face_at likewise.
struct MeshOrder {
  std::vector<int> cell_pos, cell_at, face_pos, face_at;
};

MeshOrder order_of(const amr_mesh_view& v) {
  const int n = static_cast<int>(v.n_cells), nf = static_cast<int>(v.n_faces);
  double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
  for (int k = 0; k < n; ++k) {
    x0 = std::min(x0, v.cells[k].cx);
    x1 = std::max(x1, v.cells[k].cx);
    y0 = std::min(y0, v.cells[k].cy);
    y1 = std::max(y1, v.cells[k].cy);
  }
  const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
  std::vector<std::pair<uint64_t, int>> key(n);
  for (int k = 0; k < n; ++k) {
    const uint32_t qx = static_cast<uint32_t>(std::min(65535.0, std::max(0.0, (v.cells[k].cx - x0) * sx)));
    const uint32_t qy = static_cast<uint32_t>(std::min(65535.0, std::max(0.0, (v.cells[k].cy - y0) * sy)));
    key[k] = {hilbert_index(qx, qy), k};
  }
  std::sort(key.begin(), key.end());
  MeshOrder o;
  o.cell_at.resize(n);
  o.cell_pos.resize(n);
  for (int i = 0; i < n; ++i) {
    o.cell_at[i] = key[i].second;
    o.cell_pos[key[i].second] = i;
  }
  std::vector<std::pair<int64_t, int>> fkey(nf);
  for (int q = 0; q < nf; ++q) {
    const int l = v.faces[q].left, r = v.faces[q].right;
    const int a = l >= 0 ? o.cell_pos[l] : n, b = r >= 0 ? o.cell_pos[r] : n;
    fkey[q] = {static_cast<int64_t>(std::min(a, b)) * (n + 1) + std::max(a, b), q};
  }
  std::sort(fkey.begin(), fkey.end());
  o.face_at.resize(nf);
  o.face_pos.resize(nf);
  for (int i = 0; i < nf; ++i) {
    o.face_at[i] = fkey[i].second;
    o.face_pos[fkey[i].second] = i;
  }
  return o;
}

template <class T>
struct HostArrays {
  std::vector<MeshDesc<T>> desc;
  std::vector<int> cell_block_mesh, block_tile, face_block_mesh;
  std::vector<int2> block_rec;  // the device's block tables: {tile, slot} per block_tile entry
  std::vector<T> vol, inv_vol, dt_area, wall_dist, mem_inv;
  std::vector<int> wall_first, wall_cnt, csr_off, term_off, grp_off, mem_off;
  std::vector<double> cell_cx, cell_cy;
  std::vector<int> term_src, grp_mem, mem_grp;
  std::vector<uint16_t> slot_ent;  // per cell-face entry: item index in the cell's tile << 1 | right
  std::vector<T> items_a, items_b;  // 4 + 4 per item (in the cell-face index space)
  std::vector<int> item_range;     // 2 per cell block
  std::vector<T> term_wx, term_wy, grp_w;
  std::vector<TermRec<T>> term_rec;  // the device's term records (term_src, term_wx, term_wy)
  std::vector<int> face_idx;  // host only, for the layout check: 4 per face: left, right, info, 0
  std::vector<int> wall_owner, wall_donor;
  std::vector<T> wall_nx, wall_ny, wall_len, wall_ddist;
  std::vector<double> wall_mx, wall_my;
};

enum Space { CELLS, OFFSETS, FACES, WALLS, CSR, TERMS, GROUPS, MEMBERS, TILES, TABLE };

// One uploaded array: where its host copy and device copy live, in which index space.
struct Column {
  const void* host;
  void* dev;
  size_t elem;
  size_t count;
  Space space;
};

template <class T>
class DeviceArena final : public Arena {
 public:
  ~DeviceArena() override {
    if (slab_) (void)hipFree(slab_);
    if (!external_) {
      if (pinned_) (void)hipHostFree(pinned_);
      if (pinned_frames_) (void)hipHostFree(pinned_frames_);
    }
  }

  // Lay the slab out in caller-provided device memory of `bytes` bytes instead of allocating
  // it, with host staging in pageable memory: building and using such an arena allocates no
  // device or pinned memory.  Set before upload_all; the memory is not freed.
  void use_external(void* slab, size_t bytes) {
    external_ = slab;
    external_bytes_ = bytes;
  }

  // Slots, host tables with every slot empty; no device calls.
  std::string layout(const std::vector<SlotCapacity>& caps, const ArenaFeatures& f, int tile) {
    tile_ = tile;
    features_ = f;
    int nc = 0, nf = 0, nw = 0, ncsr = 0, nt = 0, ng = 0, nm = 0;
    for (const SlotCapacity& c : caps) {
      if (c.cells < tile || c.cells % tile || c.faces % tile || c.walls % tile || c.faces < 0 || c.walls < 0 ||
          c.csr < 0 || c.terms < 0 || c.groups < 0 || c.members < 0)
        return "invalid slot capacity";
      MeshSlot s{};
      s.cell_begin = nc;
      s.face_begin = nf;
      s.wall_begin = nw;
      s.csr_begin = ncsr;
      s.term_begin = nt;
      s.grp_begin = ng;
      s.mem_begin = nm;
      s.cap = c;
      s.nv = f.sa ? 5 : 4;
      slots_.push_back(s);
      nc += c.cells;
      nf += c.faces;
      nw += c.walls;
      ncsr += c.csr;
      nt += c.terms;
      ng += c.groups;
      nm += c.members;
      if (nc >= (1 << 30) || nf >= (1 << 30) || nw >= (1 << 30) || ncsr >= (1 << 30) || nt >= (1 << 30))
        return "arena exceeds 2^30 elements in one index space";
      if (nc >= (1 << kItemRightBits) - 1) return "arena exceeds 2^27 cells (face item packing)";
    }
    if (slots_.empty()) return "arena without slots";
    totals_ = totals_of(caps, tile);
    nc_ = nc;
    nf_ = nf;
    nw_ = nw;
    HostArrays<T>& h = h_;
    h.desc.assign(slots_.size(), MeshDesc<T>{});
    for (size_t m = 0; m < slots_.size(); ++m)
      for (int b = 0; b < slots_[m].cap.cells / tile; ++b) h.cell_block_mesh.push_back(static_cast<int>(m));
    for (size_t m = 0; m < slots_.size(); ++m)
      for (int b = 0; b < slots_[m].cap.faces / tile; ++b) h.face_block_mesh.push_back(static_cast<int>(m));
    n_face_blocks_ = static_cast<int>(h.face_block_mesh.size());
    for (size_t m = 0; m < slots_.size(); ++m)
      for (int b = 0; b < slots_[m].cap.walls / tile; ++b) h.face_block_mesh.push_back(static_cast<int>(m));
    n_wall_blocks_ = static_cast<int>(h.face_block_mesh.size()) - n_face_blocks_;
    n_cell_blocks_ = static_cast<int>(h.cell_block_mesh.size());
    h.block_tile = all_block_tiles(caps, tile, &table_offsets_);
    h.block_rec.clear();
    for (const int tl : h.block_tile) h.block_rec.push_back(tl < 0 ? make_int2(-1, -1) : make_int2(tl, h.cell_block_mesh[tl]));
    if (h.face_block_mesh.empty()) h.face_block_mesh.push_back(0);
    const auto sized = [](auto& v, size_t n, auto value) { v.assign(std::max<size_t>(n, 1), value); };
    sized(h.vol, nc, T(1));
    sized(h.inv_vol, nc, T(1));
    sized(h.dt_area, nc, T(1));
    sized(h.wall_dist, nc, T(1));
    sized(h.mem_inv, nc, T(1));
    sized(h.wall_first, nc, 0);
    sized(h.wall_cnt, nc, 0);
    sized(h.cell_cx, nc, 0.0);
    sized(h.cell_cy, nc, 0.0);
    h.csr_off.assign(nc + 1, 0);
    h.term_off.assign(nc + 1, 0);
    h.grp_off.assign(nc + 1, 0);
    h.mem_off.assign(nc + 1, 0);
    sized(h.slot_ent, static_cast<size_t>(ncsr), uint16_t(0));
    sized(h.items_a, 4 * static_cast<size_t>(std::max(ncsr, 1)), T(0));
    sized(h.items_b, 4 * static_cast<size_t>(std::max(ncsr, 1)), T(0));
    h.item_range.assign(2 * static_cast<size_t>(std::max(n_cell_blocks_, 1)), 0);
    sized(h.term_src, nt, -1);
    sized(h.term_wx, nt, T(0));
    sized(h.term_wy, nt, T(0));
    TermRec<T> no_term{};
    no_term.src = -1;
    sized(h.term_rec, nt, no_term);
    sized(h.grp_mem, ng, 0);
    sized(h.grp_w, ng, T(0));
    sized(h.mem_grp, nm, 0);
    sized(h.face_idx, 4 * static_cast<size_t>(std::max(nf, 1)), 0);
    sized(h.wall_owner, nw, 0);
    sized(h.wall_donor, nw, 0);
    sized(h.wall_nx, nw, T(0));
    sized(h.wall_ny, nw, T(0));
    sized(h.wall_len, nw, T(0));
    sized(h.wall_ddist, nw, T(1));
    sized(h.wall_mx, nw, 0.0);
    sized(h.wall_my, nw, 0.0);
    sa_nu_inf_.assign(slots_.size(), 0.0);
    cell_pos_.assign(slots_.size(), {});
    for (size_t m = 0; m < slots_.size(); ++m) clear_slot(static_cast<int>(m));
    // The sentinel after the last slot.
    h.csr_off[nc] = ncsr;
    h.term_off[nc] = nt;
    h.grp_off[nc] = ng;
    h.mem_off[nc] = nm;
    return "";
  }

  // Whether mesh v fits slot m's capacity and the arena's kernel set.
  std::string fits(int m, const amr_mesh_view& v) const {
    const MeshSlot& s = slots_[m];
    const SlotCapacity need = capacity_of(v, tile_);
    if (need.cells > s.cap.cells || need.faces > s.cap.faces || need.walls > s.cap.walls || need.csr > s.cap.csr ||
        need.terms > s.cap.terms || need.groups > s.cap.groups || need.members > s.cap.members)
      return "mesh exceeds the slot capacity";
    const ArenaFeatures f = features_of(v);
    if ((f.laminar && !features_.laminar) || (f.sa && !features_.sa) || (f.small && !features_.small) ||
        (f.visc_group && !features_.visc_group))
      return "mesh needs kernels the arena was not built with";
    if (f.sa != features_.sa) return "every mesh of an arena has the same number of conserved variables";
    return "";
  }

  // Host tables of slot m for mesh `in`.
  std::string fill_slot(int m, const MeshInput& in) {
    const std::string e = fits(m, *in.view);
    if (!e.empty()) return e;
    const amr_mesh_view& v = *in.view;
    const amr_params& p = in.params;
    MeshSlot& s = slots_[m];
    HostArrays<T>& h = h_;
    const int cb = s.cell_begin, fb = s.face_begin, wb = s.wall_begin;
    const int n = static_cast<int>(v.n_cells), nf = static_cast<int>(v.n_faces), nw = static_cast<int>(v.n_walls);
    const bool laminar = v.transport.model == 1, sa = v.transport.sa_model != 0;
    const MeshOrder order = order_of(v);
    const std::vector<int>& at = order.cell_at;
    const std::vector<int>& pos_of = order.cell_pos;
    int csr = s.csr_begin, term = s.term_begin, grp = s.grp_begin, mem = s.mem_begin;
    std::vector<int> entry_face;  // mesh face of each cell-face entry of the slot, in entry order
    entry_face.reserve(static_cast<size_t>(v.face_offsets[n]));
    for (int kk = 0; kk < s.cap.cells; ++kk) {
      const int g = cb + kk;
      h.csr_off[g] = csr;
      h.term_off[g] = term;
      h.grp_off[g] = grp;
      h.mem_off[g] = mem;
      if (kk >= n) {
        pad_cell(g, wb);
        continue;
      }
      const int k = at[kk];  // the mesh's own index of the cell stored at kk
      const amr_cell& c = v.cells[k];
      h.vol[g] = static_cast<T>(c.volume);
      h.inv_vol[g] = static_cast<T>(1.0 / c.volume);
      h.dt_area[g] = static_cast<T>(v.agglomerate_small ? c.volume : (c.x1 - c.x0) * (c.y1 - c.y0));
      h.wall_dist[g] = sa ? static_cast<T>(v.wall_distance[k]) : T(1);
      h.wall_first[g] = wb + static_cast<int>(c.wall_begin);
      h.wall_cnt[g] = static_cast<int>(c.wall_count);
      h.cell_cx[g] = c.cx;
      h.cell_cy[g] = c.cy;
      for (uint64_t pos = v.face_offsets[k]; pos < v.face_offsets[k + 1]; ++pos) {
        entry_face.push_back(static_cast<int>(v.cell_faces[pos]));
        ++csr;
      }
      if (laminar)
        for (uint64_t pos = v.visc_offsets[k]; pos < v.visc_offsets[k + 1]; ++pos) {
          const amr_grad_term& t = v.visc_terms[pos];
          h.term_src[term] = t.source >= 0 ? cb + pos_of[t.source] : (t.source == -1 ? -1 : -2 - t.side);
          h.term_wx[term] = static_cast<T>(t.x);
          h.term_wy[term] = static_cast<T>(t.y);
          h.term_rec[term].src = h.term_src[term];
          h.term_rec[term].wx = h.term_wx[term];
          h.term_rec[term].wy = h.term_wy[term];
          ++term;
        }
      const uint64_t n_k = v.member_offsets[k + 1] - v.member_offsets[k];
      h.mem_inv[g] = static_cast<T>(1.0 / static_cast<double>(n_k));
      for (uint64_t pos = v.group_offsets[k]; pos < v.group_offsets[k + 1]; ++pos) {
        const uint64_t member = v.group_members[pos];
        const uint64_t n_m = v.member_offsets[member + 1] - v.member_offsets[member];
        h.grp_mem[grp] = cb + pos_of[member];
        h.grp_w[grp] = static_cast<T>(v.cells[member].volume / static_cast<double>(n_m) / v.group_volume[k]);
        ++grp;
      }
      for (uint64_t pos = v.member_offsets[k]; pos < v.member_offsets[k + 1]; ++pos)
        h.mem_grp[mem++] = cb + pos_of[v.member_groups[pos]];
    }
    for (int qq = 0; qq < s.cap.faces; ++qq) {
      const int g = fb + qq;
      if (qq >= nf) {
        pad_face(g, cb);
        continue;
      }
      const int q = order.face_at[qq];
      const amr_face& face = v.faces[q];
      int* fi = &h.face_idx[4 * static_cast<size_t>(g)];
      fi[0] = face.left >= 0 ? cb + pos_of[face.left] : -1;
      fi[1] = face.right >= 0 ? cb + pos_of[face.right] : -1;
      fi[2] = face_info(face);
      fi[3] = 0;
    }
    fill_items(m, v, order, entry_face);
    for (int j = 0; j < s.cap.walls; ++j) {
      const int g = wb + j;
      if (j >= nw) {
        pad_wall(g, cb);
        continue;
      }
      const amr_wall& w = v.walls[j];
      h.wall_owner[g] = cb + pos_of[w.owner_cell];
      h.wall_donor[g] = cb + pos_of[laminar ? v.wall_samples[j].donor : static_cast<int>(w.owner_cell)];
      h.wall_nx[g] = static_cast<T>(w.nx);
      h.wall_ny[g] = static_cast<T>(w.ny);
      h.wall_len[g] = static_cast<T>(w.length);
      h.wall_ddist[g] = static_cast<T>(laminar ? v.wall_samples[j].distance : 1.0);
      h.wall_mx[g] = (w.x0 + w.x1) * .5;
      h.wall_my[g] = (w.y0 + w.y1) * .5;
    }
    MeshDesc<T> d{};
    d.cell_begin = cb;
    d.n_cells = n;
    d.face_begin = fb;
    d.n_faces = nf;
    d.wall_begin = wb;
    d.n_walls = nw;
    d.first_cell_block = cb / tile_;
    d.n_tiles = s.cap.cells / tile_;
    d.nv = sa ? 5 : 4;
    d.laminar = laminar;
    d.sa = sa;
    d.flux = p.flux;
    d.small = v.small_cells > 0;
    d.visc_mode = (!laminar || v.transport.stabilization == 0) ? VISC_VOLUME : (d.small ? VISC_SRD_GROUP : VISC_SRD_RATE);
    d.agglomerated = v.agglomerate_small;
    d.dt_fixed = p.dt_fixed;
    d.gamma = static_cast<T>(p.gamma);
    d.cfl = static_cast<T>(p.cfl);
    d.mco = static_cast<T>(p.ausm_mco);
    d.transport.law = v.transport.law;
    d.transport.mu_ref = static_cast<T>(v.transport.mu_ref);
    d.transport.gas_constant = static_cast<T>(v.transport.gas_constant);
    d.transport.temperature_ref = static_cast<T>(v.transport.temperature_ref);
    d.transport.sutherland = static_cast<T>(v.transport.sutherland);
    d.transport.viscous_cfl = static_cast<T>(v.transport.viscous_cfl);
    d.transport.sa_nu_inf = static_cast<T>(v.transport.sa_nu_inf);
    d.transport.sa_cfl = static_cast<T>(v.transport.sa_cfl);
    for (int side = 0; side < 4; ++side) {
      const amr_boundary& b = v.boundaries[side];
      Side<T>& o = d.side[side];
      o.kind = b.kind < 0 ? p.bc[side] : b.kind;
      for (int q = 0; q < 4; ++q) o.reference[q] = static_cast<T>(b.reference[q]);
      o.total_pressure = static_cast<T>(b.total_pressure);
      o.total_temperature = static_cast<T>(b.total_temperature);
      o.gas_constant = static_cast<T>(b.gas_constant);
      o.direction[0] = static_cast<T>(b.direction[0]);
      o.direction[1] = static_cast<T>(b.direction[1]);
      o.outlet_pressure = static_cast<T>(b.outlet_pressure);
      o.wall_velocity[0] = static_cast<T>(b.wall_velocity[0]);
      o.wall_velocity[1] = static_cast<T>(b.wall_velocity[1]);
    }
    h.desc[m] = d;
    s.n_cells = n;
    s.n_faces = nf;
    s.n_walls = nw;
    s.nv = d.nv;
    sa_nu_inf_[m] = v.transport.sa_nu_inf;
    cell_pos_[m] = order.cell_pos;
    return "";
  }

  // Every index a kernel dereferences for slot m stays inside its array and inside the slot.
  std::string check_slot(int m) const {
    const HostArrays<T>& h = h_;
    const MeshSlot& s = slots_[m];
    const MeshDesc<T>& d = h.desc[m];
    if (d.n_cells != s.n_cells || d.cell_begin != s.cell_begin || d.face_begin != s.face_begin ||
        d.wall_begin != s.wall_begin || d.first_cell_block * tile_ != s.cell_begin)
      return "layout: descriptor ranges";
    if (s.n_cells == 0) return "";
    if (s.n_cells >= s.cap.cells || s.n_faces > s.cap.faces || s.n_walls > s.cap.walls) return "layout: slot overflow";
    const auto cell_ok = [&](int k) { return k >= s.cell_begin && k < s.cell_begin + s.n_cells; };
    for (int k = s.cell_begin; k < s.cell_begin + s.n_cells; ++k) {
      if (h.csr_off[k] < s.csr_begin || h.csr_off[k] > h.csr_off[k + 1] || h.csr_off[k + 1] > s.csr_begin + s.cap.csr)
        return "layout: cell-face offsets";
      const int b = k / tile_, i0 = h.item_range[2 * b], i1 = h.item_range[2 * b + 1];
      for (int j = h.csr_off[k]; j < h.csr_off[k + 1]; ++j) {
        const int e = h.slot_ent[j], i = i0 + (e >> 1);
        if (i >= i1) return "layout: cell-face slot";
        int ids[2], l, r;
        std::memcpy(ids, &h.items_a[4 * static_cast<size_t>(i)], sizeof(ids));
        l = ids[0];
        r = (ids[1] & ((1 << kItemRightBits) - 1)) - 1;
        if ((e & 1 ? r : l) != k) return "layout: cell-face orientation";
      }
      if (h.wall_cnt[k] < 0 || h.wall_first[k] < s.wall_begin || h.wall_first[k] + h.wall_cnt[k] > s.wall_begin + s.n_walls)
        return "layout: cell walls";
      if (h.term_off[k] < s.term_begin || h.term_off[k] > h.term_off[k + 1] ||
          h.term_off[k + 1] > s.term_begin + s.cap.terms)
        return "layout: gradient term offsets";
      for (int j = h.term_off[k]; j < h.term_off[k + 1]; ++j) {
        const int src = h.term_src[j];
        if (!(cell_ok(src) || (src <= -1 && src >= -5))) return "layout: gradient term source";
        const TermRec<T>& r = h.term_rec[j];
        if (r.src != src || !(r.wx == h.term_wx[j]) || !(r.wy == h.term_wy[j])) return "layout: gradient term record";
      }
      if (h.grp_off[k] < s.grp_begin || h.grp_off[k] > h.grp_off[k + 1] || h.grp_off[k + 1] > s.grp_begin + s.cap.groups ||
          h.mem_off[k] < s.mem_begin || h.mem_off[k] > h.mem_off[k + 1] || h.mem_off[k + 1] > s.mem_begin + s.cap.members)
        return "layout: SRD offsets";
      for (int j = h.grp_off[k]; j < h.grp_off[k + 1]; ++j)
        if (!cell_ok(h.grp_mem[j])) return "layout: SRD member";
      for (int j = h.mem_off[k]; j < h.mem_off[k + 1]; ++j)
        if (!cell_ok(h.mem_grp[j])) return "layout: SRD group";
      if (h.mem_off[k + 1] == h.mem_off[k]) return "layout: cell in no SRD group";
    }
    // Face items: tile ranges inside the slot's item space, in order; every item names cells of
    // the slot with a valid side, one of them in the tile; each face of the slot appears in
    // the tiles of its cells (as many items as the face_idx table implies).
    long long expected = 0;
    for (int q = s.face_begin; q < s.face_begin + s.n_faces; ++q) {
      const int* fi = &h.face_idx[4 * static_cast<size_t>(q)];
      const int tl = fi[0] >= 0 ? fi[0] / tile_ : -1, tr = fi[1] >= 0 ? fi[1] / tile_ : -1;
      expected += (tl >= 0) + (tr >= 0 && tr != tl);
    }
    int prev = s.csr_begin;
    long long found = 0;
    for (int b = s.cell_begin / tile_; b < (s.cell_begin + s.cap.cells) / tile_; ++b) {
      const int i0 = h.item_range[2 * b], i1 = h.item_range[2 * b + 1];
      if (i0 != prev || i1 < i0 || i1 > s.csr_begin + s.cap.csr) return "layout: tile item range";
      if ((i1 - i0) * 2 + 1 > 0xFFFF) return "layout: tile has too many face items";
      prev = i1;
      for (int i = i0; i < i1; ++i) {
        int ids[2];
        std::memcpy(ids, &h.items_a[4 * static_cast<size_t>(i)], sizeof(ids));
        const int l = ids[0], r = (ids[1] & ((1 << kItemRightBits) - 1)) - 1;
        const int info = (ids[1] >> kItemRightBits) & kItemInfoMask, side = (info >> 1) - 1;
        if (side < -1 || side > 3) return "layout: item info";
        // Irregular items first in the tile.
        const bool irregular = ids[1] < 0;
        if (irregular && i > i0) {
          int prev_ids[2];
          std::memcpy(prev_ids, &h.items_a[4 * static_cast<size_t>(i - 1)], sizeof(prev_ids));
          if (prev_ids[1] >= 0) return "layout: irregular item after a regular one";
        }
        if ((l != -1 && !cell_ok(l)) || (r != -1 && !cell_ok(r)) || (l < 0 && r < 0)) return "layout: item cells";
        if (side < 0 && (l < 0 || r < 0)) return "layout: interior item missing a cell";
        if (!((l >= 0 && l / tile_ == b) || (r >= 0 && r / tile_ == b))) return "layout: item outside its tile";
        ++found;
      }
    }
    if (found != expected) return "layout: face item count";
    for (int q = s.face_begin; q < s.face_begin + s.n_faces; ++q) {
      const int* fi = &h.face_idx[4 * static_cast<size_t>(q)];
      const int l = fi[0], r = fi[1], side = (fi[2] >> 1) - 1;
      if (side < -1 || side > 3) return "layout: face side";
      if (l != -1 && !cell_ok(l)) return "layout: face left";
      if (r != -1 && !cell_ok(r)) return "layout: face right";
      if (l < 0 && r < 0) return "layout: face without cell";
      if (side < 0 && (l < 0 || r < 0)) return "layout: interior face missing a cell";
    }
    for (int j = s.wall_begin; j < s.wall_begin + s.n_walls; ++j)
      if (!cell_ok(h.wall_owner[j]) || !cell_ok(h.wall_donor[j])) return "layout: wall owner/donor";
    return "";
  }

  std::string check_tables() const {
    const HostArrays<T>& h = h_;
    if (static_cast<int>(h.cell_block_mesh.size()) * tile_ != nc_) return "layout: cell block table";
    if (h.block_rec.size() != h.block_tile.size()) return "layout: block record table";
    for (size_t i = 0; i < h.block_tile.size(); ++i) {
      const int tl = h.block_tile[i];
      const bool none = tl < 0 || tl >= n_cell_blocks_;
      if (h.block_rec[i].x != (none ? -1 : tl) || h.block_rec[i].y != (none ? -1 : h.cell_block_mesh[tl]))
        return "layout: block record table";
    }
    // The whole-arena table lists every tile once; the branch tables together list every tile
    // once, each only tiles of its own slots.
    const int tables = static_cast<int>(table_offsets_.size()) - 1;
    const int branches = tables == 1 ? 1 : tables - 1;
    std::vector<int> whole(n_cell_blocks_, 0), parts(n_cell_blocks_, 0);
    for (int t = 0; t < tables; ++t) {
      const int lo = table_offsets_[t], hi = table_offsets_[t + 1];
      if ((hi - lo) % 8) return "layout: block tile table size";
      for (int i = lo; i < hi; ++i) {
        const int tl = h.block_tile[i];
        if (tl < -1 || tl >= n_cell_blocks_) return "layout: block tile table";
        if (tl < 0) continue;
        if (t == 0) {
          if (whole[tl]++) return "layout: block tile table";
        } else if (parts[tl]++ || h.cell_block_mesh[tl] % branches != t - 1) {
          return "layout: branch block tile table";
        }
      }
    }
    for (int b = 0; b < n_cell_blocks_; ++b)
      if (whole[b] != 1 || (tables > 1 && parts[b] != 1)) return "layout: block tile table misses a tile";
    for (int b = 0; b < n_cell_blocks_; ++b) {
      const MeshSlot& s = slots_[h.cell_block_mesh[b]];
      if (b * tile_ < s.cell_begin || b * tile_ >= s.cell_begin + s.cap.cells) return "layout: cell block range";
    }
    // The last stage pass commits a slot when n_tiles of its blocks have reported: every table
    // must give each slot exactly that many blocks.
    std::vector<int> per_slot(slots_.size() * tables, 0);
    for (int t = 0; t < tables; ++t)
      for (int i = table_offsets_[t]; i < table_offsets_[t + 1]; ++i)
        if (h.block_tile[i] >= 0) ++per_slot[t * slots_.size() + h.cell_block_mesh[h.block_tile[i]]];
    for (size_t m = 0; m < slots_.size(); ++m) {
      if (h.desc[m].n_tiles != slots_[m].cap.cells / tile_ || h.desc[m].n_tiles != per_slot[m])
        return "layout: slot tile count";
      for (int t = 1; t < tables; ++t)
        if (per_slot[t * slots_.size() + m] != (static_cast<int>(m) % branches == t - 1 ? h.desc[m].n_tiles : 0))
          return "layout: slot tile count in a branch table";
    }
    for (int b = 0; b < n_face_blocks_ + n_wall_blocks_; ++b) {
      const MeshSlot& s = slots_[h.face_block_mesh[b]];
      const bool face = b < n_face_blocks_;
      const int first = (face ? b : b - n_face_blocks_) * tile_;
      const int lo = face ? s.face_begin : s.wall_begin, cap = face ? s.cap.faces : s.cap.walls;
      if (first < lo || first >= lo + cap) return "layout: face/wall block range";
    }
    for (size_t m = 0; m < slots_.size(); ++m) {
      const std::string e = check_slot(static_cast<int>(m));
      if (!e.empty()) return e;
    }
    return "";
  }

  // Allocate the slab, upload every table, clear the work arrays, allocate pinned staging.
  std::string upload_all(const std::vector<amr_state_meta>& metas, hipStream_t stream) {
    const HostArrays<T>& h = h_;
    columns_.clear();
    const auto col = [&](const auto& vec, Space sp) {
      using V = typename std::decay_t<decltype(vec)>::value_type;
      columns_.push_back({vec.data(), nullptr, sizeof(V), vec.size(), sp});
    };
    col(h.desc, TABLE);
    col(h.block_rec, TABLE);
    col(h.face_block_mesh, TABLE);
    for (auto* v : {&h.vol, &h.inv_vol, &h.dt_area, &h.wall_dist, &h.mem_inv}) col(*v, CELLS);
    for (auto* v : {&h.wall_first, &h.wall_cnt}) col(*v, CELLS);
    for (auto* v : {&h.cell_cx, &h.cell_cy}) col(*v, CELLS);
    for (auto* v : {&h.csr_off, &h.term_off, &h.grp_off, &h.mem_off}) col(*v, OFFSETS);
    col(h.slot_ent, CSR);
    columns_.push_back({h.items_a.data(), nullptr, 4 * sizeof(T), h.items_a.size() / 4, CSR});
    columns_.push_back({h.items_b.data(), nullptr, 4 * sizeof(T), h.items_b.size() / 4, CSR});
    col(h.term_rec, TERMS);
    col(h.grp_mem, GROUPS);
    col(h.grp_w, GROUPS);
    col(h.mem_grp, MEMBERS);
    columns_.push_back({h.item_range.data(), nullptr, 2 * sizeof(int), h.item_range.size() / 2, TILES});
    for (auto* v : {&h.wall_owner, &h.wall_donor}) col(*v, WALLS);
    for (auto* v : {&h.wall_nx, &h.wall_ny, &h.wall_len, &h.wall_ddist}) col(*v, WALLS);
    for (auto* v : {&h.wall_mx, &h.wall_my}) col(*v, WALLS);
    size_t geo = 0;
    for (const Column& c : columns_) geo += align256(c.elem * c.count);
    const int nv = features_.sa ? 5 : 4;
    const size_t nm = slots_.size();
    const SlabPlan plan = plan_slab<T>(totals_, nv);
    if (geo != plan.geometry) return "AMR arena size bookkeeping";
    const int nc = std::max(nc_, 1), nf = std::max(nf_, 1), nw = std::max(nw_, 1);
    const size_t cplane = plan.cplane, dplane = plan.dplane, vec4 = plan.vec4, prec = plan.prec, vec8 = plan.vec8;
    const size_t mon_cell = plan.mon_cell, mon_wall = plan.mon_wall;
    slab_bytes_ = plan.total;
    geometry_bytes_ = geo;
    state_bytes_ = static_cast<uint64_t>(nv) * nc * sizeof(double);
    char* base = nullptr;
    int device = 0, lds_max = 0;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
      return "AMR device query failed";
    if (flux_lds_bytes<T>(tile_, true) > static_cast<size_t>(lds_max))
      return "AMR tile needs more LDS per block than the device provides";
    if (external_) {
      if (slab_bytes_ > external_bytes_)
        return "AMR arena needs " + std::to_string(slab_bytes_) + " bytes; the device scratch holds " +
               std::to_string(external_bytes_);
      base = static_cast<char*>(external_);
    } else {
      if (hipMalloc(reinterpret_cast<void**>(&base), slab_bytes_) != hipSuccess) return "AMR hipMalloc of the arena failed";
      slab_ = base;  // owned: freed with the arena
    }
    size_t offset = 0;
    for (Column& c : columns_) {
      c.dev = base + offset;
      offset += align256(c.elem * c.count);
    }
    ArenaPtrs<T>& a = ptrs_;
    size_t i = 0;
    const auto next = [&](auto* typed) { return reinterpret_cast<decltype(typed)>(columns_[i++].dev); };
    a.desc = next(static_cast<const MeshDesc<T>*>(nullptr));
    a.block_rec = next(static_cast<const int2*>(nullptr));
    a.face_block_mesh = next(static_cast<const int*>(nullptr));
    a.vol = next(static_cast<const T*>(nullptr));
    a.inv_vol = next(static_cast<const T*>(nullptr));
    a.dt_area = next(static_cast<const T*>(nullptr));
    a.wall_dist = next(static_cast<const T*>(nullptr));
    a.mem_inv = next(static_cast<const T*>(nullptr));
    a.wall_first = next(static_cast<const int*>(nullptr));
    a.wall_cnt = next(static_cast<const int*>(nullptr));
    a.cell_cx = next(static_cast<const double*>(nullptr));
    a.cell_cy = next(static_cast<const double*>(nullptr));
    a.csr_off = next(static_cast<const int*>(nullptr));
    a.term_off = next(static_cast<const int*>(nullptr));
    a.grp_off = next(static_cast<const int*>(nullptr));
    a.mem_off = next(static_cast<const int*>(nullptr));
    a.SLOT = next(static_cast<const uint16_t*>(nullptr));
    a.ITEM_A = next(static_cast<const T*>(nullptr));
    a.ITEM_B = next(static_cast<const T*>(nullptr));
    a.terms = next(static_cast<const TermRec<T>*>(nullptr));
    a.grp_mem = next(static_cast<const int*>(nullptr));
    a.grp_w = next(static_cast<const T*>(nullptr));
    a.mem_grp = next(static_cast<const int*>(nullptr));
    a.item_range = next(static_cast<const int2*>(nullptr));
    a.wall_owner = next(static_cast<const int*>(nullptr));
    a.wall_donor = next(static_cast<const int*>(nullptr));
    a.wall_nx = next(static_cast<const T*>(nullptr));
    a.wall_ny = next(static_cast<const T*>(nullptr));
    a.wall_len = next(static_cast<const T*>(nullptr));
    a.wall_ddist = next(static_cast<const T*>(nullptr));
    a.wall_mx = next(static_cast<const double*>(nullptr));
    a.wall_my = next(static_cast<const double*>(nullptr));
    if (i != columns_.size()) return "AMR arena column bookkeeping";
    a.ctrl = reinterpret_cast<Control*>(base + offset);
    offset += align256(nm * sizeof(Control));
    targets_ = reinterpret_cast<double*>(base + offset);
    offset += align256(nm * sizeof(double));
    maxsteps_ = reinterpret_cast<long long*>(base + offset);
    offset += align256(nm * sizeof(long long));
    frames_ = reinterpret_cast<double*>(base + offset);
    a.mon_frame = frames_;
    offset += align256(nm * 6 * sizeof(double));
    const size_t tables_end = offset;
    const auto dplanes = [&](int count) {
      double* p = reinterpret_cast<double*>(base + offset);
      offset += dplane * count;
      return p;
    };
    const auto plane = [&](size_t bytes, int count) {
      T* p = reinterpret_cast<T*>(base + offset);
      offset += bytes * count;
      return p;
    };
    a.U[0] = dplanes(nv);
    a.U[1] = dplanes(nv);
    a.U2 = dplanes(nv);
    a.QHAT = dplanes(4);
    for (int b = 0; b < 3; ++b) a.VREC[b] = plane(vec4, 1);
    for (int b = 0; b < 3; ++b) a.PREC[b] = plane(prec, 1);
    a.GREC = plane(vec8, 1);
    a.R0 = plane(cplane, nv);
    a.SRATE = plane(cplane, 1);
    a.CELL_DT = plane(cplane, 1);
    a.DRATE = plane(cplane, 1);
    a.DGROUP = plane(cplane, 1);
    a.DU = sizeof(T) == 4 ? plane(cplane, nv) : nullptr;
    a.MON_CELL = reinterpret_cast<double*>(base + offset);
    offset += mon_cell;
    a.MON_WALL = reinterpret_cast<double*>(base + offset);
    offset += mon_wall;
    if (offset != slab_bytes_) return "AMR arena layout overflow";
    a.n_meshes = static_cast<int>(nm);
    a.nc = nc;
    a.nf = nf;
    a.nw = nw;
    a.tile = tile_;
    a.flux_spread = 1;
    a.n_cell_blocks = n_cell_blocks_;
    a.n_grid_blocks = table_offsets_[1];  // the whole-arena table
    a.commit_first = 0;
    a.commit_stride = 1;
    a.n_face_blocks = n_face_blocks_;
    a.n_wall_blocks = n_wall_blocks_;
    a.any_laminar = features_.laminar;
    a.any_sa = features_.sa;
    a.any_small = features_.small;
    a.any_visc_group = features_.visc_group;
    a.commit_in_flux = features_.small ? 0 : 1;
    for (const Column& c : columns_)
      if (hipMemcpyAsync(c.dev, c.host, c.elem * c.count, hipMemcpyHostToDevice, stream) != hipSuccess)
        return "AMR geometry H2D failed";
    std::vector<Control> ctrl(nm);
    for (size_t m = 0; m < nm; ++m) ctrl[m] = initial_control(m < metas.size() ? metas[m] : amr_state_meta{0, 0, 0});
    if (hipMemcpyAsync(a.ctrl, ctrl.data(), nm * sizeof(Control), hipMemcpyHostToDevice, stream) != hipSuccess)
      return "AMR control H2D failed";
    // Zero the work arrays once so padding lanes never hold garbage (a fill kernel: an arena in
    // reused memory starts from zero as well, and the clear is a plain kernel node if captured).
    const size_t words = (slab_bytes_ - tables_end) / sizeof(uint32_t);
    if (words) {
      const unsigned blocks = static_cast<unsigned>(std::min<size_t>((words + 255) / 256, 65536));
      hipLaunchKernelGGL(k_clear_words, dim3(blocks), dim3(256), 0, stream,
                         reinterpret_cast<uint32_t*>(base + tables_end), words);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream) != hipSuccess) return "AMR work clear failed";
    size_t max_state = 1;
    for (const MeshSlot& s : slots_) max_state = std::max(max_state, static_cast<size_t>(s.nv) * s.cap.cells);
    if (external_) {
      stage_.assign(max_state, 0.0);
      stage_frames_.assign(nm * 6, 0.0);
      pinned_ = stage_.data();
      pinned_frames_ = stage_frames_.data();
    } else if (hipHostMalloc(&pinned_, max_state * sizeof(double), hipHostMallocDefault) != hipSuccess ||
               hipHostMalloc(reinterpret_cast<void**>(&pinned_frames_), nm * 6 * sizeof(double),
                             hipHostMallocDefault) != hipSuccess) {
      return "AMR pinned staging allocation failed";
    }
    return "";
  }

  // Upload the tables of slot m, its descriptor and a fresh control.
  std::string upload_slot(int m, const amr_state_meta& meta, hipStream_t stream) {
    const MeshSlot& s = slots_[m];
    for (const Column& c : columns_) {
      size_t begin = 0, count = 0;
      switch (c.space) {
        case CELLS: begin = s.cell_begin, count = s.cap.cells; break;
        case OFFSETS: begin = s.cell_begin, count = s.cap.cells; break;
        case FACES: begin = s.face_begin, count = s.cap.faces; break;
        case WALLS: begin = s.wall_begin, count = s.cap.walls; break;
        case CSR: begin = s.csr_begin, count = s.cap.csr; break;
        case TERMS: begin = s.term_begin, count = s.cap.terms; break;
        case GROUPS: begin = s.grp_begin, count = s.cap.groups; break;
        case MEMBERS: begin = s.mem_begin, count = s.cap.members; break;
        case TILES: begin = s.cell_begin / tile_, count = s.cap.cells / tile_; break;
        case TABLE: begin = 0, count = 0; break;
      }
      if (!count) continue;
      if (hipMemcpyAsync(static_cast<char*>(c.dev) + begin * c.elem, static_cast<const char*>(c.host) + begin * c.elem,
                         count * c.elem, hipMemcpyHostToDevice, stream) != hipSuccess)
        return "AMR slot H2D failed";
    }
    if (hipMemcpyAsync(const_cast<MeshDesc<T>*>(ptrs_.desc) + m, &h_.desc[m], sizeof(MeshDesc<T>), hipMemcpyHostToDevice,
                       stream) != hipSuccess)
      return "AMR descriptor H2D failed";
    const Control c = initial_control(meta);
    if (hipMemcpyAsync(ptrs_.ctrl + m, &c, sizeof(Control), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      return "AMR slot control H2D failed";
    return "";
  }

  std::string admit(int m, const MeshInput& in, hipStream_t stream) override {
    if (m < 0 || m >= static_cast<int>(slots_.size())) return "slot out of range";
    std::string e = validate_input(in);
    if (e.empty()) e = fits(m, *in.view);
    if (!e.empty()) return e;  // the slot is untouched
    e = fill_slot(m, in);
    if (e.empty()) e = check_slot(m);
    if (e.empty()) e = upload_slot(m, in.meta, stream);
    if (e.empty() && write_state(m, 0, in.u, stream) != hipSuccess) e = "AMR state H2D failed";
    if (!e.empty()) release(m, stream);
    return e;
  }

  std::string release(int m, hipStream_t stream) override {
    if (m < 0 || m >= static_cast<int>(slots_.size())) return "slot out of range";
    clear_slot(m);
    return upload_slot(m, amr_state_meta{0, 0, 0}, stream);
  }

  int precision() const override { return sizeof(T) == 8 ? 64 : 32; }
  const std::vector<MeshSlot>& slots() const override { return slots_; }
  const ArenaFeatures& features() const override { return features_; }
  uint64_t device_bytes() const override { return slab_bytes_; }
  uint64_t geometry_bytes() const override { return geometry_bytes_; }
  uint64_t state_bytes() const override { return state_bytes_; }
  int launch_trial(hipStream_t s, LaunchHook* hook) override { return hipamr::launch_trial<T>(ptrs_, s, hook); }
  int launch_residual(hipStream_t s, LaunchHook* hook) override { return hipamr::launch_residual<T>(ptrs_, s, hook); }
  int n_branches() const override { return static_cast<int>(table_offsets_.size()) == 2 ? 1 : static_cast<int>(table_offsets_.size()) - 2; }
  int launch_trial_branch(int j, hipStream_t s) override {
    if (n_branches() == 1) return hipamr::launch_trial<T>(ptrs_, s, nullptr);
    ArenaPtrs<T> a = ptrs_;  // the branch's block table; its slots commit in its own launch
    a.block_rec = ptrs_.block_rec + table_offsets_[1 + j];
    a.n_grid_blocks = table_offsets_[2 + j] - table_offsets_[1 + j];
    a.commit_first = j;
    a.commit_stride = n_branches();
    return hipamr::launch_trial<T>(a, s, nullptr);
  }
  int launch_prims(hipStream_t s) override { return hipamr::launch_prims<T>(ptrs_, s, nullptr); }
  bool set_flux_spread(int spread) override {
    if ((spread != 1 && spread != 2) || tile_ * spread > 256) return false;  // launch bound
    ptrs_.flux_spread = spread;
    return true;
  }
  int flux_spread() const override { return ptrs_.flux_spread; }
  Control* control() override { return ptrs_.ctrl; }
  double* target_array() override { return targets_; }
  long long* maxsteps_array() override { return maxsteps_; }
  hipError_t launch_monitor(hipStream_t s, const double* frames) override {
    // The previous monitor's copy has completed: read_monitor synchronises the stream.
    std::memcpy(pinned_frames_, frames, slots_.size() * 6 * sizeof(double));
    const hipError_t e =
        hipMemcpyAsync(frames_, pinned_frames_, slots_.size() * 6 * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    hipamr::launch_monitor<T>(ptrs_, s);
    return hipGetLastError();
  }
  hipError_t read_monitor(std::vector<double>* cells, std::vector<double>* walls, hipStream_t s) override {
    cells->resize(static_cast<size_t>(n_cell_blocks_) * 6);
    walls->resize(static_cast<size_t>(n_wall_blocks_) * 3);
    hipError_t e = hipMemcpyAsync(cells->data(), ptrs_.MON_CELL, cells->size() * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && !walls->empty())
      e = hipMemcpyAsync(walls->data(), ptrs_.MON_WALL, walls->size() * sizeof(double), hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
  }
  int n_meshes() const override { return static_cast<int>(slots_.size()); }
  int n_cell_blocks() const override { return n_cell_blocks_; }
  int n_face_blocks() const override { return n_face_blocks_; }
  int n_wall_blocks() const override { return n_wall_blocks_; }
  int tile() const override { return tile_; }

  hipError_t read_state(int m, int buffer, double* out, hipStream_t s) override {
    const MeshSlot& sl = slots_[m];
    const size_t count = static_cast<size_t>(sl.nv) * sl.n_cells;
    if (sl.n_cells == 0) return hipSuccess;
    const std::vector<int>& pos = cell_pos_[m];
    const size_t n = static_cast<size_t>(sl.n_cells);
    if (buffer == 2) {  // residual R0 (T planes)
      hipError_t e = hipMemcpy2DAsync(pinned_, sl.n_cells * sizeof(T), ptrs_.R0 + sl.cell_begin, nc_ * sizeof(T),
                                      sl.n_cells * sizeof(T), sl.nv, hipMemcpyDeviceToHost, s);
      if (e != hipSuccess || (e = hipStreamSynchronize(s)) != hipSuccess) return e;
      const T* p = static_cast<const T*>(pinned_);
      for (size_t v = 0; v < count / n; ++v)
        for (size_t k = 0; k < n; ++k) out[v * n + k] = static_cast<double>(p[v * n + pos[k]]);
      return hipSuccess;
    }
    hipError_t e = hipMemcpy2DAsync(pinned_, sl.n_cells * sizeof(double), ptrs_.U[buffer] + sl.cell_begin,
                                    nc_ * sizeof(double), sl.n_cells * sizeof(double), sl.nv, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess || (e = hipStreamSynchronize(s)) != hipSuccess) return e;
    const double* p = static_cast<const double*>(pinned_);
    for (size_t v = 0; v < count / n; ++v)
      for (size_t k = 0; k < n; ++k) out[v * n + k] = p[v * n + pos[k]];
    return hipSuccess;
  }

  hipError_t write_state(int m, int buffer, const double* in, hipStream_t s) override {
    const MeshSlot& sl = slots_[m];
    if (sl.n_cells == 0) return hipSuccess;
    const std::vector<int>& pos = cell_pos_[m];
    const size_t n = static_cast<size_t>(sl.n_cells);
    double* p = static_cast<double*>(pinned_);
    for (size_t v = 0; v < static_cast<size_t>(sl.nv); ++v)
      for (size_t k = 0; k < n; ++k) p[v * n + pos[k]] = in[v * n + k];
    hipError_t e = hipMemcpy2DAsync(ptrs_.U[buffer] + sl.cell_begin, nc_ * sizeof(double), pinned_,
                                    sl.n_cells * sizeof(double), sl.n_cells * sizeof(double), sl.nv,
                                    hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }

  double sa_nu_inf(int m) const override { return sa_nu_inf_[m]; }

 private:
  static Control initial_control(const amr_state_meta& meta) {
    Control c;
    std::memset(&c, 0, sizeof(c));
    c.time = meta.time;
    c.target = meta.time;
    c.steps = meta.steps;
    c.rejected = meta.rejected;
    c.finished = 1;
    c.dt_min = kInf;
    return c;
  }

  static int face_info(const amr_face& f) { return (f.axis & 1) | ((f.boundary + 1) << 1); }

  // Geometry words of face q as stored (rounded to T): b = {d_t, ex, ey, distance}; returns d_n.
  // Irregular unless d_t, ex and ey are +0 and distance is |d_n|, bit for bit.
  static T face_words(const amr_mesh_view& v, int q, bool laminar, T (&b)[4], bool* irregular) {
    const amr_face& face = v.faces[q];
    T dn = T(0);
    b[0] = b[1] = b[2] = b[3] = T(0);
    if (laminar) {
      const amr_face_geometry& geo = v.face_geometry[q];
      const bool xaxis = (face.axis & 1) == 0;
      dn = static_cast<T>(xaxis ? geo.dx : geo.dy);
      b[0] = static_cast<T>(xaxis ? geo.dy : geo.dx);
      b[1] = static_cast<T>(geo.ex);
      b[2] = static_cast<T>(geo.ey);
      b[3] = static_cast<T>(face.boundary >= 0 ? v.face_samples[q].distance : geo.distance);
    }
    const T implied[4] = {T(0), T(0), T(0), std::fabs(dn)};
    *irregular = laminar && std::memcmp(b, implied, sizeof(implied)) != 0;
    return dn;
  }

  // Face items of slot m: per tile, the faces with at least one cell in the tile (irregular
  // items first; then interior faces before boundary faces, each group in storage order), and
  // every cell-face entry's index into its tile's list.  entry_face: the mesh face of each
  // entry, in entry order.  Euler meshes read no face geometry but the length: all regular.
  void fill_items(int m, const amr_mesh_view& v, const MeshOrder& order, const std::vector<int>& entry_face) {
    const MeshSlot& s = slots_[m];
    HostArrays<T>& h = h_;
    const int cb = s.cell_begin, n = static_cast<int>(v.n_cells);
    const bool laminar = v.transport.model == 1;
    std::vector<int> slot_of(v.n_faces, -1), list;
    int item = s.csr_begin;
    for (int b = 0; b < s.cap.cells / tile_; ++b) {
      const int lo = b * tile_, hi = std::min(n, lo + tile_);
      list.clear();
      for (int kk = lo; kk < hi; ++kk)
        for (int j = h.csr_off[cb + kk]; j < h.csr_off[cb + kk + 1]; ++j) {
          const int q = entry_face[j - s.csr_begin];
          if (slot_of[q] < 0) {
            slot_of[q] = 0;
            list.push_back(q);
          }
        }
      for (const int q : list) {
        T words[4];
        bool irregular;
        face_words(v, q, laminar, words, &irregular);
        slot_of[q] = irregular ? 1 : 0;  // sort key only; replaced by the index below
      }
      std::sort(list.begin(), list.end(), [&](int x, int y) {
        if (slot_of[x] != slot_of[y]) return slot_of[x] > slot_of[y];
        const bool bx = v.faces[x].boundary >= 0, by = v.faces[y].boundary >= 0;
        return bx != by ? by : order.face_pos[x] < order.face_pos[y];
      });
      const int tb = cb / tile_ + b;
      h.item_range[2 * tb] = item;
      for (size_t i = 0; i < list.size(); ++i) {
        const int q = list[i];
        slot_of[q] = static_cast<int>(i);
        const amr_face& face = v.faces[q];
        T* wa = &h.items_a[4 * static_cast<size_t>(item)];
        T* wb = &h.items_b[4 * static_cast<size_t>(item++)];
        std::fill(wa, wa + 4, T(0));
        bool irregular;
        const T dn = face_words(v, q, laminar, *reinterpret_cast<T(*)[4]>(wb), &irregular);
        const unsigned right_info = static_cast<unsigned>(face.right >= 0 ? cb + order.cell_pos[face.right] + 1 : 0) |
                                    (static_cast<unsigned>(face_info(face)) << kItemRightBits) |
                                    (irregular ? 1u << 31 : 0u);
        const int ids[2] = {face.left >= 0 ? cb + order.cell_pos[face.left] : -1, static_cast<int>(right_info)};
        std::memcpy(wa, ids, sizeof(ids));
        wa[2] = static_cast<T>(face.length);
        wa[3] = dn;
        if (!irregular) std::fill(wb, wb + 4, T(0));
      }
      h.item_range[2 * tb + 1] = item;
      for (int kk = lo; kk < hi; ++kk) {
        const int k = order.cell_at[kk];
        for (int j = h.csr_off[cb + kk]; j < h.csr_off[cb + kk + 1]; ++j) {
          const int q = entry_face[j - s.csr_begin];
          const int right = v.faces[q].right == k ? 1 : 0;
          h.slot_ent[j] = static_cast<uint16_t>((slot_of[q] << 1) | right);
        }
      }
      for (const int q : list) slot_of[q] = -1;
    }
  }

  void pad_cell(int g, int wb) {
    HostArrays<T>& h = h_;
    h.vol[g] = h.inv_vol[g] = h.dt_area[g] = h.wall_dist[g] = h.mem_inv[g] = T(1);
    h.wall_first[g] = wb;
    h.wall_cnt[g] = 0;
    h.cell_cx[g] = h.cell_cy[g] = 0.0;
  }
  void pad_face(int g, int cb) {
    HostArrays<T>& h = h_;
    int* fi = &h.face_idx[4 * static_cast<size_t>(g)];
    fi[0] = fi[1] = cb;
    fi[2] = fi[3] = 0;
  }
  void pad_wall(int g, int cb) {
    HostArrays<T>& h = h_;
    h.wall_owner[g] = h.wall_donor[g] = cb;
    h.wall_nx[g] = h.wall_ny[g] = h.wall_len[g] = T(0);
    h.wall_ddist[g] = T(1);
    h.wall_mx[g] = h.wall_my[g] = 0.0;
  }
  void clear_slot(int m) {
    MeshSlot& s = slots_[m];
    HostArrays<T>& h = h_;
    for (int k = 0; k < s.cap.cells; ++k) {
      const int g = s.cell_begin + k;
      h.csr_off[g] = s.csr_begin;
      h.term_off[g] = s.term_begin;
      h.grp_off[g] = s.grp_begin;
      h.mem_off[g] = s.mem_begin;
      pad_cell(g, s.wall_begin);
    }
    for (int q = 0; q < s.cap.faces; ++q) pad_face(s.face_begin + q, s.cell_begin);
    for (int b = s.cell_begin / tile_; b < (s.cell_begin + s.cap.cells) / tile_; ++b)
      h.item_range[2 * b] = h.item_range[2 * b + 1] = s.csr_begin;
    for (int j = 0; j < s.cap.walls; ++j) pad_wall(s.wall_begin + j, s.cell_begin);
    MeshDesc<T> d{};
    d.cell_begin = s.cell_begin;
    d.face_begin = s.face_begin;
    d.wall_begin = s.wall_begin;
    d.first_cell_block = s.cell_begin / tile_;
    d.n_tiles = s.cap.cells / tile_;
    d.nv = features_.sa ? 5 : 4;
    d.gamma = T(1.4);
    h.desc[m] = d;
    s.n_cells = s.n_faces = s.n_walls = 0;
    sa_nu_inf_[m] = 0.0;
    cell_pos_[m].clear();
  }

  HostArrays<T> h_;
  std::vector<std::vector<int>> cell_pos_;  // per slot: storage position of each mesh cell
  std::vector<Column> columns_;
  ArenaPtrs<T> ptrs_{};
  std::vector<MeshSlot> slots_;
  std::vector<double> sa_nu_inf_;
  ArenaFeatures features_{};
  char* slab_ = nullptr;  // owned slab (nullptr for an external one)
  void* external_ = nullptr;
  size_t external_bytes_ = 0;
  std::vector<double> stage_, stage_frames_;  // pageable staging of an external arena
  void* pinned_ = nullptr;
  double* targets_ = nullptr;
  long long* maxsteps_ = nullptr;
  double* frames_ = nullptr;
  double* pinned_frames_ = nullptr;
  Totals totals_{};
  uint64_t slab_bytes_ = 0, geometry_bytes_ = 0, state_bytes_ = 0;
  int nc_ = 0, nf_ = 0, nw_ = 0, tile_ = 256;
  int n_cell_blocks_ = 0, n_face_blocks_ = 0, n_wall_blocks_ = 0;
  std::vector<int> table_offsets_;  // block tables in block_tile / block_rec: whole arena, then each branch
};

template <class T>
std::unique_ptr<Arena> build_arena(const std::vector<SlotCapacity>& caps, const ArenaFeatures& f, int tile,
                                   const std::vector<MeshInput>* inputs, hipStream_t stream, std::string* error,
                                   void* slab = nullptr, size_t slab_bytes = 0) {
  auto a = std::make_unique<DeviceArena<T>>();
  if (slab) a->use_external(slab, slab_bytes);
  std::string e = a->layout(caps, f, tile);
  std::vector<amr_state_meta> metas;
  if (e.empty() && inputs)
    for (size_t m = 0; m < inputs->size() && e.empty(); ++m) {
      e = a->fill_slot(static_cast<int>(m), (*inputs)[m]);
      metas.push_back((*inputs)[m].meta);
    }
  if (e.empty()) e = a->check_tables();
  if (e.empty()) e = a->upload_all(metas, stream);
  if (e.empty() && inputs)
    for (size_t m = 0; m < inputs->size() && e.empty(); ++m)
      if (a->write_state(static_cast<int>(m), 0, (*inputs)[m].u, stream) != hipSuccess) e = "AMR initial state H2D failed";
  if (!e.empty()) {
    *error = e;
    return nullptr;
  }
  return a;
}

}  // namespace

SlotCapacity capacity_of(const amr_mesh_view& v, int tile) {
  SlotCapacity c{};
  c.cells = round_up(static_cast<int>(v.n_cells) + 1, tile);
  c.faces = round_up(static_cast<int>(v.n_faces), tile);
  c.walls = round_up(static_cast<int>(v.n_walls), tile);
  c.csr = static_cast<int>(v.face_offsets[v.n_cells]);
  c.terms = v.visc_offsets ? static_cast<int>(v.visc_offsets[v.n_cells]) : 0;
  c.groups = static_cast<int>(v.group_offsets[v.n_cells]);
  c.members = static_cast<int>(v.member_offsets[v.n_cells]);
  return c;
}

ArenaFeatures features_of(const amr_mesh_view& v) {
  ArenaFeatures f{};
  f.laminar = v.transport.model == 1;
  f.sa = v.transport.sa_model != 0;
  f.small = v.small_cells > 0;
  f.visc_group = f.laminar && v.transport.stabilization != 0 && f.small;
  return f;
}

std::string validate_input(const MeshInput& in) {
  const amr_mesh_view* v = in.view;
  const amr_params& p = in.params;
  if (!v || !in.u) return "invalid AMR GPU configuration";
  if (!tables_ready(*v) || v->n_cells == 0 || v->n_faces == 0 || !std::isfinite(in.meta.time) || in.meta.time < 0 ||
      p.limiter != 3 || p.time_integrator != 0 || !std::isfinite(p.gamma) || !(p.gamma > 1) || !(p.cfl > 0) ||
      p.cfl > .5 || !std::isfinite(p.dt_fixed) || p.dt_fixed < 0 || !std::isfinite(p.ausm_mco) || p.ausm_mco <= 0 ||
      (p.flux != 0 && p.flux != 1))
    return "invalid AMR GPU configuration";
  for (int side = 0; side < 4; ++side)
    if (!side_valid(v->boundaries[side], p.bc[side], side)) return "invalid AMR GPU boundary configuration";
  if (v->n_cells > (1u << 29) || v->n_faces > (1u << 29) || v->n_walls > (1u << 29))
    return "AMR mesh exceeds 2^29 elements";
  return "";
}

std::unique_ptr<Arena> make_arena(const std::vector<MeshInput>& inputs, int precision, int tile, hipStream_t stream,
                                  std::string* error, void* slab, size_t slab_bytes) {
  if (tile != 64 && tile != 128 && tile != 256) {
    *error = "tile must be 64, 128 or 256";
    return nullptr;
  }
  std::vector<SlotCapacity> caps;
  ArenaFeatures f{};
  for (const MeshInput& in : inputs) {
    const std::string e = validate_input(in);
    if (!e.empty()) {
      *error = e;
      return nullptr;
    }
    caps.push_back(capacity_of(*in.view, tile));
    const ArenaFeatures g = features_of(*in.view);
    f.laminar |= g.laminar;
    f.sa |= g.sa;
    f.small |= g.small;
    f.visc_group |= g.visc_group;
  }
  if (precision == 64) return build_arena<double>(caps, f, tile, &inputs, stream, error, slab, slab_bytes);
  if (precision == 32) return build_arena<float>(caps, f, tile, &inputs, stream, error, slab, slab_bytes);
  *error = "precision must be 64 or 32";
  return nullptr;
}

std::unique_ptr<Arena> make_slot_arena(const std::vector<SlotCapacity>& caps, const ArenaFeatures& features,
                                       int precision, int tile, hipStream_t stream, std::string* error) {
  if (tile != 64 && tile != 128 && tile != 256) {
    *error = "tile must be 64, 128 or 256";
    return nullptr;
  }
  if (precision == 64) return build_arena<double>(caps, features, tile, nullptr, stream, error);
  if (precision == 32) return build_arena<float>(caps, features, tile, nullptr, stream, error);
  *error = "precision must be 64 or 32";
  return nullptr;
}

uint64_t arena_bytes(const std::vector<SlotCapacity>& caps, const ArenaFeatures& features, int precision, int tile) {
  const Totals t = totals_of(caps, tile);
  const int nv = features.sa ? 5 : 4;
  return precision == 32 ? plan_slab<float>(t, nv).total : plan_slab<double>(t, nv).total;
}

std::string check_inputs_layout(const std::vector<MeshInput>& inputs, int tile) {
  std::vector<SlotCapacity> caps;
  ArenaFeatures f{1, 0, 1, 1};
  for (const MeshInput& in : inputs) {
    caps.push_back(capacity_of(*in.view, tile));
    f.sa = in.view->transport.sa_model != 0;
  }
  DeviceArena<double> arena;
  std::string e = arena.layout(caps, f, tile);
  for (size_t m = 0; m < inputs.size() && e.empty(); ++m) e = arena.fill_slot(static_cast<int>(m), inputs[m]);
  return e.empty() ? arena.check_tables() : e;
}

void sum_monitor(const Arena& arena, int m, const std::vector<double>& cells, const std::vector<double>& walls,
                 double* out) {
  const MeshSlot& s = arena.slots()[m];
  const int tile = arena.tile();
  long double acc[9] = {};
  for (int b = s.cell_begin / tile; b < (s.cell_begin + s.cap.cells) / tile; ++b)
    for (int i = 0; i < 6; ++i) acc[i] += cells[6 * b + i];
  for (int b = s.wall_begin / tile; b < (s.wall_begin + s.cap.walls) / tile; ++b)
    for (int i = 0; i < 3; ++i) acc[6 + i] += walls[3 * b + i];
  for (int i = 0; i < 9; ++i) out[i] = static_cast<double>(acc[i]);
}

}  // namespace hipamr

extern "C" int amr_hip_check_layout(const amr_mesh_view* views, const amr_params* params, int32_t n, int32_t tile,
                                    char* err, size_t err_len) {
  std::string message;
  if (!views || !params || n < 1 || (tile != 64 && tile != 128 && tile != 256)) message = "invalid layout check arguments";
  std::vector<std::vector<double>> dummy(message.empty() ? n : 0);
  std::vector<hipamr::MeshInput> inputs;
  for (int32_t m = 0; message.empty() && m < n; ++m) {
    hipamr::MeshInput in{};
    in.view = views + m;
    in.params = params[m];
    dummy[m].assign((views[m].transport.sa_model ? 5 : 4) * views[m].n_cells, 1.0);
    in.u = dummy[m].data();
    in.meta = amr_state_meta{0, 0, 0};
    message = hipamr::validate_input(in);
    inputs.push_back(in);
  }
  if (message.empty()) message = hipamr::check_inputs_layout(inputs, tile);
  if (message.empty()) return 0;
  if (err && err_len) {
    const size_t k = message.size() < err_len - 1 ? message.size() : err_len - 1;
    std::memcpy(err, message.data(), k);
    err[k] = 0;
  }
  return 1;
}
