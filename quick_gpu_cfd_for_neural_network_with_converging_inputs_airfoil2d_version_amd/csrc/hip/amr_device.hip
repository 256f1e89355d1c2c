// amr_device.hip -- arena layout, one-allocation upload and state transfer (see amr_device.h).
//
// Layout: mesh m gets tile-aligned global ranges [cell_begin, cell_begin + n_cells) etc.; the
// padding tail of a range belongs to no cell (all CSR offsets stay monotone there).  Every
// array lives in one hipMalloc'ed slab (256-byte aligned sub-allocations), so an arena costs a
// single allocation and frees in one call.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "amr_device.h"

namespace hipamr {
namespace {

constexpr int BC_TRANSMISSIVE = 1, BC_REFLECT = 2, BC_FARFIELD = 3, BC_TOTAL_INLET = 4, BC_PRESSURE_OUTLET = 5,
              BC_NOSLIP = 6;

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

template <class T>
struct HostArrays {
  std::vector<MeshDesc<T>> desc;
  std::vector<int> cell_block_mesh, face_block_mesh;
  std::vector<T> vol, inv_vol, dt_area, wall_dist;
  std::vector<int> csr_off, csr_face, wall_first, wall_cnt, term_off, term_src;
  std::vector<T> term_wx, term_wy;
  std::vector<int> grp_off, grp_mem, mem_off, mem_grp;
  std::vector<T> grp_w, mem_inv;
  std::vector<int> face_l, face_r, face_info;
  std::vector<T> face_len, face_dx, face_dy, face_ex, face_ey, face_dist;
  std::vector<int> wall_owner, wall_donor;
  std::vector<T> wall_nx, wall_ny, wall_len, wall_ddist;
};

template <class T>
class DeviceArena final : public Arena {
 public:
  ~DeviceArena() override {
    if (slab_) (void)hipFree(slab_);
    if (pinned_) (void)hipHostFree(pinned_);
  }

  std::string build(const std::vector<MeshInput>& in, int tile, hipStream_t stream) {
    HostArrays<T> h;
    std::string message = layout(in, tile, &h);
    if (message.empty()) message = check_layout(h);
    if (!message.empty()) return message;
    return upload(h, in, stream);
  }

  // Host-side arrays of the arena (no device calls).
  std::string layout(const std::vector<MeshInput>& in, int tile, HostArrays<T>* hp) {
    HostArrays<T>& h = *hp;
    tile_ = tile;
    int nc = 0, nf = 0, nw = 0;
    for (const MeshInput& m : in) {
      MeshSlot s{};
      s.cell_begin = nc;
      s.n_cells = static_cast<int>(m.view->n_cells);
      s.face_begin = nf;
      s.n_faces = static_cast<int>(m.view->n_faces);
      s.wall_begin = nw;
      s.n_walls = static_cast<int>(m.view->n_walls);
      s.nv = m.view->transport.sa_model ? 5 : 4;
      nc += round_up(s.n_cells, tile);
      nf += round_up(s.n_faces, tile);
      nw += round_up(s.n_walls, tile);
      slots_.push_back(s);
    }
    nc_ = nc;
    nf_ = nf;
    nw_ = nw;
    h.vol.assign(nc, T(1));
    h.inv_vol.assign(nc, T(1));
    h.dt_area.assign(nc, T(1));
    h.wall_dist.assign(nc, T(1));
    h.wall_first.assign(nc, 0);
    h.wall_cnt.assign(nc, 0);
    h.mem_inv.assign(nc, T(1));
    h.face_l.assign(nf, 0);
    h.face_r.assign(nf, 0);
    h.face_info.assign(nf, 0);
    h.face_len.assign(nf, T(0));
    h.face_dx.assign(nf, T(0));
    h.face_dy.assign(nf, T(0));
    h.face_ex.assign(nf, T(0));
    h.face_ey.assign(nf, T(0));
    h.face_dist.assign(nf, T(1));
    h.wall_owner.assign(nw, 0);
    h.wall_donor.assign(nw, 0);
    h.wall_nx.assign(nw, T(0));
    h.wall_ny.assign(nw, T(0));
    h.wall_len.assign(nw, T(0));
    h.wall_ddist.assign(nw, T(1));
    h.csr_off.push_back(0);
    h.term_off.push_back(0);
    h.grp_off.push_back(0);
    h.mem_off.push_back(0);
    for (size_t mi = 0; mi < in.size(); ++mi) {
      const amr_mesh_view& v = *in[mi].view;
      const amr_params& p = in[mi].params;
      const MeshSlot& s = slots_[mi];
      const int cb = s.cell_begin, fb = s.face_begin, wb = s.wall_begin;
      const bool laminar = v.transport.model == 1, sa = v.transport.sa_model != 0;
      for (int k = 0; k < round_up(s.n_cells, tile); ++k) {
        const int g = cb + k;
        if (k < s.n_cells) {
          const amr_cell& c = v.cells[k];
          h.vol[g] = static_cast<T>(c.volume);
          h.inv_vol[g] = static_cast<T>(1.0 / c.volume);
          h.dt_area[g] = static_cast<T>(v.agglomerate_small ? c.volume : (c.x1 - c.x0) * (c.y1 - c.y0));
          if (sa) h.wall_dist[g] = static_cast<T>(v.wall_distance[k]);
          h.wall_first[g] = wb + static_cast<int>(c.wall_begin);
          h.wall_cnt[g] = static_cast<int>(c.wall_count);
          for (uint64_t pos = v.face_offsets[k]; pos < v.face_offsets[k + 1]; ++pos) {
            const int q = static_cast<int>(v.cell_faces[pos]);
            h.csr_face.push_back(v.faces[q].left == k ? fb + q : ~(fb + q));
          }
          if (laminar)
            for (uint64_t pos = v.visc_offsets[k]; pos < v.visc_offsets[k + 1]; ++pos) {
              const amr_grad_term& t = v.visc_terms[pos];
              h.term_src.push_back(t.source >= 0 ? cb + t.source : (t.source == -1 ? -1 : -2 - t.side));
              h.term_wx.push_back(static_cast<T>(t.x));
              h.term_wy.push_back(static_cast<T>(t.y));
            }
          const uint64_t n_k = v.member_offsets[k + 1] - v.member_offsets[k];
          h.mem_inv[g] = static_cast<T>(1.0 / static_cast<double>(n_k));
          for (uint64_t pos = v.group_offsets[k]; pos < v.group_offsets[k + 1]; ++pos) {
            const uint64_t member = v.group_members[pos];
            const uint64_t n_m = v.member_offsets[member + 1] - v.member_offsets[member];
            h.grp_mem.push_back(cb + static_cast<int>(member));
            h.grp_w.push_back(static_cast<T>(v.cells[member].volume / static_cast<double>(n_m) / v.group_volume[k]));
          }
          for (uint64_t pos = v.member_offsets[k]; pos < v.member_offsets[k + 1]; ++pos)
            h.mem_grp.push_back(cb + static_cast<int>(v.member_groups[pos]));
        }
        h.csr_off.push_back(static_cast<int>(h.csr_face.size()));
        h.term_off.push_back(static_cast<int>(h.term_src.size()));
        h.grp_off.push_back(static_cast<int>(h.grp_mem.size()));
        h.mem_off.push_back(static_cast<int>(h.mem_grp.size()));
      }
      for (int q = 0; q < s.n_faces; ++q) {
        const amr_face& f = v.faces[q];
        const int g = fb + q;
        h.face_l[g] = f.left >= 0 ? cb + f.left : -1;
        h.face_r[g] = f.right >= 0 ? cb + f.right : -1;
        h.face_info[g] = (f.axis & 1) | ((f.boundary + 1) << 1);
        h.face_len[g] = static_cast<T>(f.length);
        if (laminar) {
          const amr_face_geometry& geo = v.face_geometry[q];
          h.face_dx[g] = static_cast<T>(geo.dx);
          h.face_dy[g] = static_cast<T>(geo.dy);
          h.face_ex[g] = static_cast<T>(geo.ex);
          h.face_ey[g] = static_cast<T>(geo.ey);
          h.face_dist[g] = static_cast<T>(f.boundary >= 0 ? v.face_samples[q].distance : geo.distance);
        }
      }
      for (int j = 0; j < s.n_walls; ++j) {
        const amr_wall& w = v.walls[j];
        const int g = wb + j;
        h.wall_owner[g] = cb + static_cast<int>(w.owner_cell);
        h.wall_donor[g] = laminar ? cb + v.wall_samples[j].donor : cb + static_cast<int>(w.owner_cell);
        h.wall_nx[g] = static_cast<T>(w.nx);
        h.wall_ny[g] = static_cast<T>(w.ny);
        h.wall_len[g] = static_cast<T>(w.length);
        h.wall_ddist[g] = static_cast<T>(laminar ? v.wall_samples[j].distance : 1.0);
      }
      MeshDesc<T> d{};
      d.cell_begin = cb;
      d.n_cells = s.n_cells;
      d.face_begin = fb;
      d.n_faces = s.n_faces;
      d.wall_begin = wb;
      d.n_walls = s.n_walls;
      d.first_cell_block = cb / tile;
      d.nv = s.nv;
      d.laminar = laminar;
      d.sa = sa;
      d.flux = p.flux;
      d.small = v.small_cells > 0;
      d.visc_mode = (!laminar || v.transport.stabilization == 0) ? VISC_VOLUME
                    : (d.small ? VISC_SRD_GROUP : VISC_SRD_RATE);
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
      h.desc.push_back(d);
      for (int b = 0; b < round_up(s.n_cells, tile) / tile; ++b) h.cell_block_mesh.push_back(static_cast<int>(mi));
      any_laminar_ |= laminar;
      any_sa_ |= sa;
      any_small_ |= d.small != 0;
      any_visc_group_ |= d.visc_mode == VISC_SRD_GROUP;
    }
    for (size_t mi = 0; mi < in.size(); ++mi)
      for (int b = 0; b < round_up(slots_[mi].n_faces, tile) / tile; ++b) h.face_block_mesh.push_back(static_cast<int>(mi));
    n_face_blocks_ = static_cast<int>(h.face_block_mesh.size());
    for (size_t mi = 0; mi < in.size(); ++mi)
      for (int b = 0; b < round_up(slots_[mi].n_walls, tile) / tile; ++b) h.face_block_mesh.push_back(static_cast<int>(mi));
    n_wall_blocks_ = static_cast<int>(h.face_block_mesh.size()) - n_face_blocks_;
    n_cell_blocks_ = static_cast<int>(h.cell_block_mesh.size());
    // Non-empty placeholders keep every device pointer valid.
    for (auto* v : {&h.csr_face, &h.term_src, &h.grp_mem, &h.mem_grp})
      if (v->empty()) v->push_back(0);
    for (auto* v : {&h.term_wx, &h.term_wy, &h.grp_w})
      if (v->empty()) v->push_back(T(0));
    if (h.face_block_mesh.empty()) h.face_block_mesh.push_back(0);
    return "";
  }

  // Every index a kernel dereferences stays inside its array and inside its own mesh's range
  // (the bounds the launch grid and the kernels assume).  Empty string when consistent.
  std::string check_layout(const HostArrays<T>& h) const {
    const int n_meshes = static_cast<int>(slots_.size());
    if (static_cast<int>(h.desc.size()) != n_meshes) return "layout: descriptor count";
    if (static_cast<int>(h.cell_block_mesh.size()) * tile_ != nc_) return "layout: cell block table";
    if (n_face_blocks_ * tile_ != nf_ || n_wall_blocks_ * tile_ != nw_) return "layout: face/wall block table";
    if (h.csr_off.size() != static_cast<size_t>(nc_) + 1 || h.term_off.size() != static_cast<size_t>(nc_) + 1 ||
        h.grp_off.size() != static_cast<size_t>(nc_) + 1 || h.mem_off.size() != static_cast<size_t>(nc_) + 1)
      return "layout: offset array sizes";
    for (int b = 0; b < static_cast<int>(h.cell_block_mesh.size()); ++b) {
      const int m = h.cell_block_mesh[b];
      if (m < 0 || m >= n_meshes) return "layout: cell block mesh id";
      const MeshSlot& s = slots_[m];
      if (b * tile_ < s.cell_begin || b * tile_ >= s.cell_begin + std::max(s.n_cells, 1)) return "layout: cell block range";
    }
    for (int b = 0; b < n_face_blocks_ + n_wall_blocks_; ++b) {
      const int m = h.face_block_mesh[b];
      if (m < 0 || m >= n_meshes) return "layout: face block mesh id";
      const MeshSlot& s = slots_[m];
      const int first = b < n_face_blocks_ ? b * tile_ : (b - n_face_blocks_) * tile_;
      const int lo = b < n_face_blocks_ ? s.face_begin : s.wall_begin;
      const int n = b < n_face_blocks_ ? s.n_faces : s.n_walls;
      if (first < lo || first >= lo + n) return "layout: face/wall block range";
    }
    for (int m = 0; m < n_meshes; ++m) {
      const MeshSlot& s = slots_[m];
      const MeshDesc<T>& d = h.desc[m];
      if (d.cell_begin != s.cell_begin || d.n_cells != s.n_cells || d.face_begin != s.face_begin ||
          d.n_faces != s.n_faces || d.wall_begin != s.wall_begin || d.n_walls != s.n_walls ||
          d.first_cell_block * tile_ != s.cell_begin)
        return "layout: descriptor ranges";
      const auto cell_ok = [&](int k) { return k >= s.cell_begin && k < s.cell_begin + s.n_cells; };
      for (int k = s.cell_begin; k < s.cell_begin + s.n_cells; ++k) {
        if (h.csr_off[k] > h.csr_off[k + 1] || h.csr_off[k + 1] > static_cast<int>(h.csr_face.size()))
          return "layout: cell-face offsets";
        for (int j = h.csr_off[k]; j < h.csr_off[k + 1]; ++j) {
          const int sq = h.csr_face[j], q = sq >= 0 ? sq : ~sq;
          if (q < s.face_begin || q >= s.face_begin + s.n_faces) return "layout: cell-face id";
          if ((sq >= 0 ? h.face_l[q] : h.face_r[q]) != k) return "layout: cell-face orientation";
        }
        if (h.wall_cnt[k] < 0 || h.wall_first[k] < s.wall_begin || h.wall_first[k] + h.wall_cnt[k] > s.wall_begin + s.n_walls)
          return "layout: cell walls";
        if (h.term_off[k] > h.term_off[k + 1] || h.term_off[k + 1] > static_cast<int>(h.term_src.size()) ||
            h.term_off[k + 1] > static_cast<int>(h.term_wx.size()))
          return "layout: gradient term offsets";
        for (int j = h.term_off[k]; j < h.term_off[k + 1]; ++j) {
          const int src = h.term_src[j];
          if (!(cell_ok(src) || (src <= -1 && src >= -5))) return "layout: gradient term source";
        }
        if (h.grp_off[k + 1] > static_cast<int>(h.grp_mem.size()) || h.mem_off[k + 1] > static_cast<int>(h.mem_grp.size()))
          return "layout: SRD offsets";
        for (int j = h.grp_off[k]; j < h.grp_off[k + 1]; ++j)
          if (!cell_ok(h.grp_mem[j])) return "layout: SRD member";
        for (int j = h.mem_off[k]; j < h.mem_off[k + 1]; ++j)
          if (!cell_ok(h.mem_grp[j])) return "layout: SRD group";
        if (h.mem_off[k + 1] == h.mem_off[k]) return "layout: cell in no SRD group";
      }
      for (int q = s.face_begin; q < s.face_begin + s.n_faces; ++q) {
        const int l = h.face_l[q], r = h.face_r[q], side = (h.face_info[q] >> 1) - 1;
        if (side < -1 || side > 3) return "layout: face side";
        if (l != -1 && !cell_ok(l)) return "layout: face left";
        if (r != -1 && !cell_ok(r)) return "layout: face right";
        if (l < 0 && r < 0) return "layout: face without cell";
        if (side < 0 && (l < 0 || r < 0)) return "layout: interior face missing a cell";
      }
      for (int j = s.wall_begin; j < s.wall_begin + s.n_walls; ++j)
        if (!cell_ok(h.wall_owner[j]) || !cell_ok(h.wall_donor[j])) return "layout: wall owner/donor";
    }
    return "";
  }

  int precision() const override { return sizeof(T) == 8 ? 64 : 32; }
  const std::vector<MeshSlot>& slots() const override { return slots_; }
  uint64_t device_bytes() const override { return slab_bytes_; }
  uint64_t geometry_bytes() const override { return geometry_bytes_; }
  uint64_t state_bytes() const override { return state_bytes_; }
  int launch_trial(hipStream_t s) override { return hipamr::launch_trial<T>(ptrs_, s); }
  int launch_residual(hipStream_t s) override { return hipamr::launch_residual<T>(ptrs_, s); }
  Control* control() override { return ptrs_.ctrl; }
  int n_meshes() const override { return static_cast<int>(slots_.size()); }
  int n_cell_blocks() const override { return n_cell_blocks_; }
  int n_face_blocks() const override { return n_face_blocks_; }
  int n_wall_blocks() const override { return n_wall_blocks_; }
  int tile() const override { return tile_; }

  hipError_t read_state(int m, int buffer, double* out, hipStream_t s) override {
    const MeshSlot& sl = slots_[m];
    const T* src = buffer == 2 ? ptrs_.R0 : ptrs_.U[buffer];
    hipError_t e = hipMemcpy2DAsync(pinned_, sl.n_cells * sizeof(T), src + sl.cell_begin, nc_ * sizeof(T),
                                    sl.n_cells * sizeof(T), sl.nv, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    const T* p = static_cast<const T*>(pinned_);
    for (size_t i = 0; i < static_cast<size_t>(sl.nv) * sl.n_cells; ++i) out[i] = static_cast<double>(p[i]);
    return hipSuccess;
  }

  hipError_t write_state(int m, int buffer, const double* in, hipStream_t s) override {
    const MeshSlot& sl = slots_[m];
    T* p = static_cast<T*>(pinned_);
    for (size_t i = 0; i < static_cast<size_t>(sl.nv) * sl.n_cells; ++i) p[i] = static_cast<T>(in[i]);
    hipError_t e = hipMemcpy2DAsync(ptrs_.U[buffer] + sl.cell_begin, nc_ * sizeof(T), pinned_, sl.n_cells * sizeof(T),
                                    sl.n_cells * sizeof(T), sl.nv, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }

  std::string error_;

 private:
  template <class V>
  static size_t bytes_of(const std::vector<V>& v) {
    return (v.size() * sizeof(V) + 255) / 256 * 256;
  }

  std::string upload(const HostArrays<T>& h, const std::vector<MeshInput>& in, hipStream_t stream) {
    const int nc = nc_, nf = std::max(nf_, 1), nw = std::max(nw_, 1);
    const size_t cplane = (static_cast<size_t>(nc) * sizeof(T) + 255) / 256 * 256;
    const size_t fplane = (static_cast<size_t>(nf) * sizeof(T) + 255) / 256 * 256;
    const size_t wplane = (static_cast<size_t>(nw) * sizeof(T) + 255) / 256 * 256;
    // Geometry + tables.
    size_t geo = 0;
    geo += bytes_of(h.desc) + bytes_of(h.cell_block_mesh) + bytes_of(h.face_block_mesh);
    for (auto* v : {&h.vol, &h.inv_vol, &h.dt_area, &h.wall_dist, &h.term_wx, &h.term_wy, &h.grp_w, &h.mem_inv,
                    &h.face_len, &h.face_dx, &h.face_dy, &h.face_ex, &h.face_ey, &h.face_dist, &h.wall_nx, &h.wall_ny,
                    &h.wall_len, &h.wall_ddist})
      geo += bytes_of(*v);
    for (auto* v : {&h.csr_off, &h.csr_face, &h.wall_first, &h.wall_cnt, &h.term_off, &h.term_src, &h.grp_off,
                    &h.grp_mem, &h.mem_off, &h.mem_grp, &h.face_l, &h.face_r, &h.face_info, &h.wall_owner,
                    &h.wall_donor})
      geo += bytes_of(*v);
    const int nv = any_sa_ ? 5 : 4;
    const size_t ctrl_bytes = (in.size() * sizeof(Control) + 255) / 256 * 256;
    // Work arrays: U[2], U2, R0 (nv planes each), W (4), PHI, GRAD (4), MU, EDDY, SGRAD (2),
    // SDIFF, SSRC, SRATE, QHAT (4), CELL_DT, DRATE, DGROUP; faces FLUX (4), SPEED, FCOEF, SFLUX,
    // SFCOEF; walls WFLUX (2), WCOEF, SWFLUX, SWCOEF.
    const size_t cell_planes = 4 * nv + 4 + 1 + 4 + 2 + 2 + 3 + 4 + 3;
    const size_t work = cell_planes * cplane + 8 * fplane + 5 * wplane;
    slab_bytes_ = geo + ctrl_bytes + work;
    geometry_bytes_ = geo;
    state_bytes_ = static_cast<uint64_t>(nv) * nc * sizeof(T);
    char* base = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&base), slab_bytes_) != hipSuccess) return "AMR hipMalloc of the arena failed";
    slab_ = base;
    size_t offset = 0;
    std::vector<char> staging(geo + ctrl_bytes, 0);
    const auto put = [&](const auto& vec) {
      using V = typename std::decay_t<decltype(vec)>::value_type;
      V* dst = reinterpret_cast<V*>(base + offset);
      std::memcpy(staging.data() + offset, vec.data(), vec.size() * sizeof(V));
      offset += bytes_of(vec);
      return dst;
    };
    ArenaPtrs<T>& a = ptrs_;
    a.desc = put(h.desc);
    a.cell_block_mesh = put(h.cell_block_mesh);
    a.face_block_mesh = put(h.face_block_mesh);
    a.vol = put(h.vol);
    a.inv_vol = put(h.inv_vol);
    a.dt_area = put(h.dt_area);
    a.wall_dist = put(h.wall_dist);
    a.term_wx = put(h.term_wx);
    a.term_wy = put(h.term_wy);
    a.grp_w = put(h.grp_w);
    a.mem_inv = put(h.mem_inv);
    a.face_len = put(h.face_len);
    a.face_dx = put(h.face_dx);
    a.face_dy = put(h.face_dy);
    a.face_ex = put(h.face_ex);
    a.face_ey = put(h.face_ey);
    a.face_dist = put(h.face_dist);
    a.wall_nx = put(h.wall_nx);
    a.wall_ny = put(h.wall_ny);
    a.wall_len = put(h.wall_len);
    a.wall_ddist = put(h.wall_ddist);
    a.csr_off = put(h.csr_off);
    a.csr_face = put(h.csr_face);
    a.wall_first = put(h.wall_first);
    a.wall_cnt = put(h.wall_cnt);
    a.term_off = put(h.term_off);
    a.term_src = put(h.term_src);
    a.grp_off = put(h.grp_off);
    a.grp_mem = put(h.grp_mem);
    a.mem_off = put(h.mem_off);
    a.mem_grp = put(h.mem_grp);
    a.face_l = put(h.face_l);
    a.face_r = put(h.face_r);
    a.face_info = put(h.face_info);
    a.wall_owner = put(h.wall_owner);
    a.wall_donor = put(h.wall_donor);
    // Controls: initial time and counters of every mesh; finished until the first target.
    std::vector<Control> ctrl(in.size());
    for (size_t m = 0; m < in.size(); ++m) {
      Control& c = ctrl[m];
      std::memset(&c, 0, sizeof(c));
      c.time = in[m].meta.time;
      c.target = in[m].meta.time;
      c.steps = in[m].meta.steps;
      c.rejected = in[m].meta.rejected;
      c.finished = 1;
      c.dt_min = 0x7ff0000000000000ull;
    }
    a.ctrl = put(ctrl);
    const size_t uploaded = offset;
    const auto plane = [&](size_t bytes, int count) {
      T* p = reinterpret_cast<T*>(base + offset);
      offset += bytes * count;
      return p;
    };
    a.U[0] = plane(cplane, nv);
    a.U[1] = plane(cplane, nv);
    a.U2 = plane(cplane, nv);
    a.R0 = plane(cplane, nv);
    // Kernels index planes as v * nc + k; nc is a multiple of the tile, so nc * sizeof(T)
    // is already a multiple of 256 and equals the reserved plane size.
    a.W = plane(cplane, 4);
    a.PHI = plane(cplane, 1);
    a.GRAD = plane(cplane, 4);
    a.MU = plane(cplane, 1);
    a.EDDY = plane(cplane, 1);
    a.SGRAD = plane(cplane, 2);
    a.SDIFF = plane(cplane, 1);
    a.SSRC = plane(cplane, 1);
    a.SRATE = plane(cplane, 1);
    a.QHAT = plane(cplane, 4);
    a.CELL_DT = plane(cplane, 1);
    a.DRATE = plane(cplane, 1);
    a.DGROUP = plane(cplane, 1);
    a.FLUX = plane(fplane, 4);
    a.SPEED = plane(fplane, 1);
    a.FCOEF = plane(fplane, 1);
    a.SFLUX = plane(fplane, 1);
    a.SFCOEF = plane(fplane, 1);
    a.WFLUX = plane(wplane, 2);
    a.WCOEF = plane(wplane, 1);
    a.SWFLUX = plane(wplane, 1);
    a.SWCOEF = plane(wplane, 1);
    if (offset > slab_bytes_) return "AMR arena layout overflow";
    a.n_meshes = static_cast<int>(in.size());
    a.nc = nc;
    a.nf = nf;
    a.nw = nw;
    a.tile = tile_;
    a.n_cell_blocks = n_cell_blocks_;
    a.n_face_blocks = n_face_blocks_;
    a.n_wall_blocks = n_wall_blocks_;
    a.any_laminar = any_laminar_;
    a.any_sa = any_sa_;
    a.any_small = any_small_;
    a.any_visc_group = any_visc_group_;
    if (hipMemcpyAsync(base, staging.data(), uploaded, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      return "AMR geometry H2D failed";
    // Zero the work arrays once so padding lanes never hold garbage.
    if (hipMemsetAsync(base + uploaded, 0, slab_bytes_ - uploaded, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      return "AMR work clear failed";
    size_t max_state = 0;
    for (const MeshSlot& s : slots_) max_state = std::max(max_state, static_cast<size_t>(s.nv) * s.n_cells);
    if (hipHostMalloc(&pinned_, std::max<size_t>(max_state, 1) * sizeof(T), hipHostMallocDefault) != hipSuccess)
      return "AMR pinned staging allocation failed";
    pinned_bytes_ = std::max<size_t>(max_state, 1) * sizeof(T);
    for (size_t m = 0; m < in.size(); ++m)
      if (write_state(static_cast<int>(m), 0, in[m].u, stream) != hipSuccess) return "AMR initial state H2D failed";
    return "";
  }

  ArenaPtrs<T> ptrs_{};
  std::vector<MeshSlot> slots_;
  char* slab_ = nullptr;
  void* pinned_ = nullptr;
  size_t pinned_bytes_ = 0;
  uint64_t slab_bytes_ = 0, geometry_bytes_ = 0, state_bytes_ = 0;
  int nc_ = 0, nf_ = 0, nw_ = 0, tile_ = 256;
  int n_cell_blocks_ = 0, n_face_blocks_ = 0, n_wall_blocks_ = 0;
  int any_laminar_ = 0, any_sa_ = 0, any_small_ = 0, any_visc_group_ = 0;
};

}  // namespace

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
  if (v->n_cells > (1u << 30) || v->n_faces > (1u << 30) || v->n_walls > (1u << 30))
    return "AMR mesh exceeds 2^30 elements";
  return "";
}

std::unique_ptr<Arena> make_arena(const std::vector<MeshInput>& inputs, int precision, int tile, hipStream_t stream,
                                  std::string* error) {
  if (tile != 64 && tile != 128 && tile != 256) {
    *error = "tile must be 64, 128 or 256";
    return nullptr;
  }
  size_t nc = 0, nf = 0, nw = 0;
  for (const MeshInput& m : inputs) {
    nc += round_up(static_cast<int>(m.view->n_cells), tile);
    nf += round_up(static_cast<int>(m.view->n_faces), tile);
    nw += round_up(static_cast<int>(m.view->n_walls), tile);
  }
  if (nc >= (1u << 30) || nf >= (1u << 30) || nw >= (1u << 30)) {
    *error = "arena exceeds 2^30 cells, faces or walls";
    return nullptr;
  }
  std::string message;
  if (precision == 64) {
    auto a = std::make_unique<DeviceArena<double>>();
    message = a->build(inputs, tile, stream);
    if (message.empty()) return a;
  } else if (precision == 32) {
    auto a = std::make_unique<DeviceArena<float>>();
    message = a->build(inputs, tile, stream);
    if (message.empty()) return a;
  } else {
    message = "precision must be 64 or 32";
  }
  *error = message;
  return nullptr;
}

std::string check_inputs_layout(const std::vector<MeshInput>& inputs, int tile) {
  DeviceArena<double> arena;
  HostArrays<double> h;
  std::string message = arena.layout(inputs, tile, &h);
  return message.empty() ? arena.check_layout(h) : message;
}

}  // namespace hipamr

extern "C" int amr_hip_check_layout(const amr_mesh_view* views, const amr_params* params, int32_t n, int32_t tile,
                                    char* err, size_t err_len) {
  std::string message;
  if (!views || !params || n < 1 || (tile != 64 && tile != 128 && tile != 256)) message = "invalid layout check arguments";
  std::vector<std::vector<double>> dummy;
  std::vector<hipamr::MeshInput> inputs;
  for (int32_t m = 0; message.empty() && m < n; ++m) {
    hipamr::MeshInput in{};
    in.view = views + m;
    in.params = params[m];
    dummy.emplace_back((views[m].transport.sa_model ? 5 : 4) * views[m].n_cells, 1.0);
    in.u = dummy.back().data();
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
