// api_caller.cpp -- a caller of the C++ backend API (include/static_amr2d_gpu.h) written the way
// the reference's solver drives it (src/quick_airfoil.cpp:61-97): a cfd::StaticAmrMesh2D and a
// cfd::StaticAmrState2D held by the caller, static_amr_gpu_create, advance to a target with and
// without a wall budget, download, stats, destroy, and the error paths.
//
// usage: api_caller DIR
// DIR holds the mesh as raw arrays (<name>.bin, the bytes of the reference records), params.bin
// (Params2D), u0.bin and scalars.txt ("small_cells agglomerate t1 t2").  Writes u.bin (state
// after advancing to t2) and result.txt (one "key value..." line per observation).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "static_amr2d_gpu.h"

namespace {

template <class T>
bool read_array(const std::string& path, std::vector<T>* out, bool required = true) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    if (required) std::cerr << "missing " << path << "\n";
    return !required;
  }
  const std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (bytes.size() % sizeof(T)) {
    std::cerr << "bad size " << path << "\n";
    return false;
  }
  out->resize(bytes.size() / sizeof(T));
  if (!bytes.empty()) std::memcpy(static_cast<void*>(out->data()), bytes.data(), bytes.size());
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::cerr << "usage: api_caller DIR\n";
    return 64;
  }
  const std::string dir = argv[1];
  const auto at = [&](const char* name) { return dir + "/" + name + ".bin"; };
  cfd::StaticAmrMesh2D mesh;
  std::vector<cfd::AmrBoundary2D> sides;
  std::vector<cfd::AmrTransport2D> transport;
  std::vector<cfd::Params2D> params;
  cfd::StaticAmrState2D state;
  bool ok = read_array(at("cells"), &mesh.cells) && read_array(at("faces"), &mesh.faces) &&
            read_array(at("walls"), &mesh.walls) && read_array(at("face_offsets"), &mesh.face_offsets) &&
            read_array(at("cell_faces"), &mesh.cell_faces) && read_array(at("group_offsets"), &mesh.group_offsets) &&
            read_array(at("group_members"), &mesh.group_members) &&
            read_array(at("member_offsets"), &mesh.member_offsets) &&
            read_array(at("member_groups"), &mesh.member_groups) && read_array(at("group_volume"), &mesh.group_volume) &&
            read_array(at("visc_offsets"), &mesh.viscous.offsets, false) &&
            read_array(at("visc_terms"), &mesh.viscous.terms, false) &&
            read_array(at("wall_samples"), &mesh.viscous.wall_samples, false) &&
            read_array(at("face_samples"), &mesh.viscous.face_samples, false) &&
            read_array(at("face_geometry"), &mesh.viscous.face_geometry, false) &&
            read_array(at("wall_distance"), &mesh.viscous.wall_distance, false) && read_array(at("boundaries"), &sides) &&
            read_array(at("transport"), &transport) && read_array(at("params"), &params) && read_array(at("u0"), &state.u);
  std::ifstream scalars(dir + "/scalars.txt");
  size_t small = 0;
  int agglomerate = 0;
  double t1 = 0, t2 = 0;
  ok = ok && sides.size() == 4 && transport.size() == 1 && params.size() == 1 &&
       static_cast<bool>(scalars >> small >> agglomerate >> t1 >> t2);
  if (!ok) {
    std::cerr << "api_caller: bad input in " << dir << "\n";
    return 65;
  }
  for (int s = 0; s < 4; ++s) mesh.boundaries[s] = sides[s];
  mesh.transport = transport[0];
  mesh.small_cells = small;
  mesh.options.agglomerate_small = agglomerate != 0;
  const cfd::Params2D p = params[0];
  const cfd::StaticAmrState2D initial = state;
  std::ofstream result(dir + "/result.txt");
  result.precision(17);

  // 1. The solver's loop: create, advance to t1, download, advance to t2, download.
  cfd::AmrGpuOptions options;
  options.fused = true;
  cfd::StaticAmrGpu* raw = nullptr;
  std::string error;
  int status = cfd::static_amr_gpu_create(mesh, p, state, options, &raw, &error);
  result << "create " << status << "\n";
  if (status) {
    std::cerr << error << "\n";
    return 1;
  }
  for (const double target : {t1, t2}) {
    status = cfd::static_amr_gpu_advance(raw, target, 50000000, &state, &error);
    const int down = cfd::static_amr_gpu_download(raw, &state, &error);
    result << "advance " << status << " " << down << " " << state.time << " " << state.steps << " " << state.rejected
           << "\n";
  }
  const cfd::AmrGpuStats stats = cfd::static_amr_gpu_stats(raw);
  result << "stats " << stats.status_polls << " " << stats.full_downloads << " " << stats.state_d2h_bytes << " "
         << stats.allocated_bytes << " " << stats.graph_nodes_per_trial << "\n";
  cfd::static_amr_gpu_destroy(raw);
  {
    std::ofstream u(dir + "/u.bin", std::ios::binary);
    u.write(reinterpret_cast<const char*>(state.u.data()), static_cast<std::streamsize>(state.u.size() * sizeof(double)));
  }

  // 2. Wall budget: a positive budget that has expired at the first poll returns 2 with a
  //    resumable state; the next call without a budget reaches the target.
  state = initial;
  status = cfd::static_amr_gpu_create(mesh, p, state, options, &raw, &error);
  int budget_status = cfd::static_amr_gpu_advance(raw, t1, 50000000, &state, &error, 1e-9);
  const long budget_steps = state.steps;
  const double budget_time = state.time;
  const int resume_status = cfd::static_amr_gpu_advance(raw, t1, 50000000, &state, &error);
  cfd::StaticAmrState2D resumed;
  const int down = cfd::static_amr_gpu_download(raw, &resumed, &error);
  result << "budget " << status << " " << budget_status << " " << budget_time << " " << budget_steps << " "
         << resume_status << " " << down << " " << resumed.time << " " << resumed.steps << " "
         << (resumed.u.size() == initial.u.size() ? 1 : 0) << "\n";
  {
    std::ofstream u(dir + "/u_resumed.bin", std::ios::binary);
    u.write(reinterpret_cast<const char*>(resumed.u.data()), static_cast<std::streamsize>(resumed.u.size() * sizeof(double)));
  }
  // An advance to a target behind the session's time is invalid (status 1, message).
  error.clear();
  status = cfd::static_amr_gpu_advance(raw, 0.5 * t1, 50000000, &state, &error);
  result << "behind " << status << " " << error << "\n";
  cfd::static_amr_gpu_destroy(raw);

  // 3. Invalid configurations: CFL above 0.5, batch_trials out of range, a state of the wrong
  //    size.  Each returns 1 with a message and no session.
  cfd::Params2D bad = p;
  bad.cfl = 0.7;
  raw = nullptr;
  error.clear();
  status = cfd::static_amr_gpu_create(mesh, bad, initial, options, &raw, &error);
  result << "cfl " << status << " " << (raw == nullptr ? 1 : 0) << " " << error << "\n";
  cfd::AmrGpuOptions many = options;
  many.batch_trials = 0;
  error.clear();
  status = cfd::static_amr_gpu_create(mesh, p, initial, many, &raw, &error);
  result << "trials " << status << " " << (raw == nullptr ? 1 : 0) << " " << error << "\n";
  cfd::StaticAmrState2D shorter = initial;
  shorter.u.pop_back();
  error.clear();
  status = cfd::static_amr_gpu_create(mesh, p, shorter, options, &raw, &error);
  result << "size " << status << " " << (raw == nullptr ? 1 : 0) << " " << error << "\n";

  // 4. Null sessions: stats are zero, destroy is a no-op, advance and download fail.
  const cfd::AmrGpuStats zero = cfd::static_amr_gpu_stats(nullptr);
  cfd::static_amr_gpu_destroy(nullptr);
  error.clear();
  const int a = cfd::static_amr_gpu_advance(nullptr, t1, 1, &state, &error);
  std::string e2;
  const int d = cfd::static_amr_gpu_download(nullptr, &state, &e2);
  result << "null " << zero.status_polls + zero.allocated_bytes << " " << a << " " << d << " " << error << "|" << e2
         << "\n";
  return 0;
}
