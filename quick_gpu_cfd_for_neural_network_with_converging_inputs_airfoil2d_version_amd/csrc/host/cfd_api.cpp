// cfd_api.cpp -- the reference's C++ backend API (static_amr2d_gpu.h) over the C ABI.
#include <cstring>

#include "amr_hip.h"
#include "mesher.h"
#include "static_amr2d_gpu.h"
#include "views.h"

namespace cfd {

struct StaticAmrGpu {
  amr_session* session = nullptr;
  size_t values = 0;
};

static_assert(sizeof(AmrGpuStats) == sizeof(amr_gpu_stats), "AmrGpuStats layout");

int static_amr_gpu_create(const StaticAmrMesh2D& m, const Params2D& p, const StaticAmrState2D& initial,
                          const AmrGpuOptions& options, StaticAmrGpu** result, std::string* error) {
  if (!result) return 1;
  *result = nullptr;
  const size_t nv = m.transport.sa_model ? 5 : 4;
  if (!hipamr::viscous_ready(m) || m.cells.empty() || m.faces.empty() || initial.u.size() != nv * m.cells.size()) {
    if (error) *error = "invalid AMR GPU configuration";
    return 1;
  }
  amr_mesh_view view;
  hipamr::fill_view(m, &view);
  const amr_params params = hipamr::params_to_c(p);
  const amr_state_meta meta{initial.time, initial.steps, initial.rejected};
  amr_session_options o{};
  o.batch_trials = options.batch_trials;
  o.graph = options.graph ? 1 : 0;
  o.event_timing = options.event_timing ? 1 : 0;
  o.precision = 64;
  char message[512] = {0};
  amr_session* s = nullptr;
  if (amr_hip_create(&view, &params, initial.u.data(), initial.u.size(), &meta, &o, &s, message, sizeof(message))) {
    if (error) *error = message;
    return 1;
  }
  *result = new StaticAmrGpu{s, initial.u.size()};
  return 0;
}

int static_amr_gpu_advance(StaticAmrGpu* g, double tend, long maxsteps, StaticAmrState2D* state, std::string* error,
                           double wall_budget_seconds) {
  if (!g || !state) {
    if (error) *error = "invalid AMR GPU advance";
    return 1;
  }
  amr_state_meta meta{state->time, state->steps, state->rejected};
  char message[512] = {0};
  const int status = amr_hip_advance(g->session, tend, maxsteps, wall_budget_seconds, &meta, message, sizeof(message));
  state->time = meta.time;
  state->steps = meta.steps;
  state->rejected = meta.rejected;
  if (status == 1 && error) *error = message;
  return status;
}

int static_amr_gpu_download(StaticAmrGpu* g, StaticAmrState2D* state, std::string* error) {
  if (!g || !state) {
    if (error) *error = "invalid AMR GPU download";
    return 1;
  }
  state->u.resize(g->values);
  amr_state_meta meta{};
  char message[512] = {0};
  if (amr_hip_download(g->session, state->u.data(), state->u.size(), &meta, message, sizeof(message))) {
    if (error) *error = message;
    return 1;
  }
  state->time = meta.time;
  state->steps = meta.steps;
  state->rejected = meta.rejected;
  return 0;
}

AmrGpuStats static_amr_gpu_stats(const StaticAmrGpu* g) {
  AmrGpuStats out;
  if (!g) return out;
  amr_gpu_stats raw{};
  amr_hip_stats(g->session, &raw);
  std::memcpy(static_cast<void*>(&out), &raw, sizeof(out));
  return out;
}

void static_amr_gpu_destroy(StaticAmrGpu* g) {
  if (!g) return;
  amr_hip_destroy(g->session);
  delete g;
}

}  // namespace cfd
