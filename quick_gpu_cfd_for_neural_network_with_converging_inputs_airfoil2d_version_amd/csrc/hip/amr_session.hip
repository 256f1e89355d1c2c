// amr_session.hip -- one device-resident solve behind the C ABI (amr_hip_* in amr_hip.h).
//
// Semantics follow the reference backend (src/static_amr2d_cuda.cu:219-366): create validates,
// uploads and captures one trial as a graph; advance replays it batch_trials times per host
// poll of the 104-byte control record until the target time is reached (0), a fatal code is
// raised (1, last accepted state retained) or the wall budget expires (2); download copies
// the accepted state; stats reports the same counters as AmrGpuStats.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "amr_device.h"

using hipamr::Arena;
using hipamr::Control;
using hipamr::MeshInput;

struct amr_session {
  std::unique_ptr<Arena> arena;
  amr_session_options options{};
  amr_gpu_stats stats{};
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t start = nullptr, end = nullptr;
  Control* host_control = nullptr;  // pinned
  int device = 0;
  uint64_t n_cells = 0, nv = 0;

  ~amr_session() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (start) (void)hipEventDestroy(start);
    if (end) (void)hipEventDestroy(end);
    arena.reset();
    if (host_control) (void)hipHostFree(host_control);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

using Clock = std::chrono::steady_clock;

int fail(char* err, size_t len, const std::string& message) {
  if (err && len) {
    const size_t k = message.size() < len - 1 ? message.size() : len - 1;
    std::memcpy(err, message.data(), k);
    err[k] = 0;
  }
  return 1;
}

bool checked(hipError_t status, char* err, size_t len, const char* operation) {
  if (status == hipSuccess) return true;
  fail(err, len, std::string(operation) + ": " + hipGetErrorString(status));
  return false;
}

void to_meta(const Control& c, amr_state_meta* meta) {
  if (!meta) return;
  meta->time = c.time;
  meta->steps = c.steps;
  meta->rejected = c.rejected;
}

// Synchronous copy ordered on the session stream.
hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t stream) {
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, stream);
  return e != hipSuccess ? e : hipStreamSynchronize(stream);
}

int launch_one(amr_session* s) {
  if (s->options.graph && s->exec) return hipGraphLaunch(s->exec, s->stream) == hipSuccess ? 0 : 1;
  s->arena->launch_trial(s->stream);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace

extern "C" {

int amr_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int amr_hip_create(const amr_mesh_view* mesh, const amr_params* params, const double* u, uint64_t u_len,
                   const amr_state_meta* initial, const amr_session_options* options, amr_session** out, char* err,
                   size_t err_len) {
  if (!out) return fail(err, err_len, "null session pointer");
  *out = nullptr;
  if (!mesh || !params || !u || !options) return fail(err, err_len, "invalid AMR GPU configuration");
  const auto t0 = Clock::now();
  MeshInput in{};
  in.view = mesh;
  in.params = *params;
  in.u = u;
  in.meta = initial ? *initial : amr_state_meta{0, 0, 0};
  const uint64_t nv = mesh->transport.sa_model ? 5 : 4;
  std::string message = hipamr::validate_input(in);
  if (message.empty() && u_len != nv * mesh->n_cells) message = "invalid AMR GPU configuration";
  if (message.empty() && (options->batch_trials < 1 || options->batch_trials > 1024))
    message = "invalid AMR GPU configuration";
  if (message.empty() && options->precision != 64 && options->precision != 32) message = "precision must be 64 or 32";
  if (!message.empty()) return fail(err, err_len, message);
  if (amr_hip_device_count() < 1) return fail(err, err_len, "no HIP device available for the AMR backend");
  std::unique_ptr<amr_session> s(new amr_session);
  s->options = *options;
  s->n_cells = mesh->n_cells;
  s->nv = nv;
  if (!checked(hipGetDevice(&s->device), err, err_len, "AMR device")) return 1;
  if (!checked(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), err, err_len, "AMR stream")) return 1;
  if (options->event_timing && (!checked(hipEventCreate(&s->start), err, err_len, "AMR timing start") ||
                                !checked(hipEventCreate(&s->end), err, err_len, "AMR timing end")))
    return 1;
  // A small mesh is latency-bound: 64-wide tiles spread it over more compute units.
  const int tile = mesh->n_cells < 200000 ? 64 : 256;
  s->arena = hipamr::make_arena({in}, options->precision, tile, s->stream, &message);
  if (!s->arena) return fail(err, err_len, message);
  if (!checked(hipHostMalloc(reinterpret_cast<void**>(&s->host_control), sizeof(Control), hipHostMallocDefault), err,
               err_len, "AMR pinned control"))
    return 1;
  if (!checked(hipMemcpyAsync(s->host_control, s->arena->control(), sizeof(Control), hipMemcpyDeviceToHost, s->stream),
               err, err_len, "AMR control D2H") ||
      !checked(hipStreamSynchronize(s->stream), err, err_len, "AMR setup sync"))
    return 1;
  s->stats.allocated_bytes = s->arena->device_bytes();
  s->stats.pinned_bytes = sizeof(Control) + nv * mesh->n_cells * (options->precision == 64 ? 8 : 4);
  s->stats.geometry_h2d_bytes = s->arena->geometry_bytes();
  s->stats.state_h2d_bytes = nv * mesh->n_cells * (options->precision == 64 ? 8 : 4);
  s->stats.control_h2d_bytes = sizeof(Control);
  if (options->graph) {
    if (!checked(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), err, err_len, "AMR graph begin"))
      return 1;
    s->arena->launch_trial(s->stream);
    if (!checked(hipStreamEndCapture(s->stream, &s->graph), err, err_len, "AMR graph end") ||
        !checked(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0), err, err_len, "AMR graph instantiate"))
      return 1;
    size_t nodes = 0;
    if (!checked(hipGraphGetNodes(s->graph, nullptr, &nodes), err, err_len, "AMR graph node count")) return 1;
    s->stats.graph_nodes_per_trial = nodes;
  }
  s->stats.setup_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
  *out = s.release();
  return 0;
}

int amr_hip_advance(amr_session* s, double tend, int64_t maxsteps, double wall_budget_seconds, amr_state_meta* state,
                    char* err, size_t err_len) {
  if (!s || !state || !std::isfinite(tend) || tend < s->host_control->time || maxsteps < 1 ||
      !std::isfinite(wall_budget_seconds) || wall_budget_seconds < 0)
    return fail(err, err_len, "invalid AMR GPU advance");
  const auto t0 = Clock::now();
  auto last = t0;
  bool expired = false;
  hipamr::launch_target(s->arena->control(), 1, tend, maxsteps, nullptr, nullptr, s->stream);
  for (;;) {
    if (s->options.event_timing && !checked(hipEventRecord(s->start, s->stream), err, err_len, "AMR timing record start"))
      return 1;
    for (int trial = 0; trial < s->options.batch_trials; ++trial)
      if (launch_one(s)) return fail(err, err_len, "AMR trial launch failed");
    if (s->options.event_timing && !checked(hipEventRecord(s->end, s->stream), err, err_len, "AMR timing record end"))
      return 1;
    if (!checked(hipGetLastError(), err, err_len, "AMR trial kernels") ||
        !checked(hipMemcpyAsync(s->host_control, s->arena->control(), sizeof(Control), hipMemcpyDeviceToHost, s->stream),
                 err, err_len, "AMR control D2H") ||
        !checked(hipStreamSynchronize(s->stream), err, err_len, "AMR batch sync"))
      return 1;
    if (s->options.event_timing) {
      float ms = 0;
      if (!checked(hipEventElapsedTime(&ms, s->start, s->end), err, err_len, "AMR timing elapsed")) return 1;
      s->stats.advance_event_seconds += ms * .001;
      ++s->stats.event_batches;
    }
    s->stats.control_d2h_bytes += sizeof(Control);
    ++s->stats.status_polls;
    to_meta(*s->host_control, state);
    const auto now = Clock::now();
    if (std::chrono::duration<double>(now - last).count() >= 2) {
      std::fprintf(stderr, "amr-hip: t=%.9g steps=%lld rejected=%lld polls=%llu\n", state->time,
                   static_cast<long long>(state->steps), static_cast<long long>(state->rejected),
                   static_cast<unsigned long long>(s->stats.status_polls));
      last = now;
    }
    if (s->host_control->fatal || s->host_control->finished) break;
    if (wall_budget_seconds > 0 && std::chrono::duration<double>(now - t0).count() >= wall_budget_seconds) {
      expired = true;
      break;
    }
  }
  s->stats.advance_seconds += std::chrono::duration<double>(Clock::now() - t0).count();
  if (s->host_control->fatal)
    return fail(err, err_len,
                "AMR GPU stopped with status " + std::to_string(s->host_control->fatal) + "; last accepted state retained");
  return expired ? 2 : 0;
}

int amr_hip_download(amr_session* s, double* u, uint64_t u_len, amr_state_meta* state, char* err, size_t err_len) {
  if (!s || !u || u_len != s->nv * s->n_cells) return fail(err, err_len, "invalid AMR GPU download");
  const auto t0 = Clock::now();
  if (!checked(hipMemcpyAsync(s->host_control, s->arena->control(), sizeof(Control), hipMemcpyDeviceToHost, s->stream),
               err, err_len, "AMR control D2H") ||
      !checked(hipStreamSynchronize(s->stream), err, err_len, "AMR download sync") ||
      !checked(s->arena->read_state(0, s->host_control->cur, u, s->stream), err, err_len, "AMR state D2H"))
    return 1;
  to_meta(*s->host_control, state);
  s->stats.state_d2h_bytes += s->nv * s->n_cells * (s->options.precision == 64 ? 8 : 4);
  ++s->stats.full_downloads;
  s->stats.download_seconds += std::chrono::duration<double>(Clock::now() - t0).count();
  return 0;
}

void amr_hip_stats(const amr_session* s, amr_gpu_stats* out) {
  if (!out) return;
  if (!s) {
    std::memset(out, 0, sizeof(*out));
    return;
  }
  *out = s->stats;
}

void amr_hip_destroy(amr_session* s) { delete s; }

int amr_hip_rhs(amr_session* s, const double* u, uint64_t u_len, double* r, char* err, size_t err_len) {
  if (!s || !u || !r || u_len != s->nv * s->n_cells) return fail(err, err_len, "invalid AMR GPU residual request");
  // Save the accepted state and control, evaluate L(u) in place, restore.
  Control saved{};
  std::vector<double> keep(u_len);
  if (!checked(copy_sync(&saved, s->arena->control(), sizeof(Control), hipMemcpyDeviceToHost, s->stream), err, err_len,
               "AMR control D2H") ||
      !checked(s->arena->read_state(0, saved.cur, keep.data(), s->stream), err, err_len, "AMR state D2H"))
    return 1;
  Control eval = saved;
  eval.active = 1;
  eval.fatal = 0;
  eval.bad = 0;
  eval.w_valid = 0;
  eval.r0_valid = 0;
  eval.dt_min = 0x7ff0000000000000ull;
  Control after{};
  if (!checked(copy_sync(s->arena->control(), &eval, sizeof(Control), hipMemcpyHostToDevice, s->stream), err, err_len,
               "AMR control H2D") ||
      !checked(s->arena->write_state(0, saved.cur, u, s->stream), err, err_len, "AMR state H2D"))
    return 1;
  s->arena->launch_residual(s->stream);
  if (!checked(hipGetLastError(), err, err_len, "AMR residual kernels") ||
      !checked(hipStreamSynchronize(s->stream), err, err_len, "AMR residual sync") ||
      !checked(copy_sync(&after, s->arena->control(), sizeof(Control), hipMemcpyDeviceToHost, s->stream), err, err_len,
               "AMR control D2H"))
    return 1;
  const bool ok = !after.bad && !after.fatal;
  if (ok && !checked(s->arena->read_state(0, 2, r, s->stream), err, err_len, "AMR residual D2H")) return 1;
  saved.w_valid = 0;
  saved.r0_valid = 0;
  saved.dt_min = 0x7ff0000000000000ull;
  if (!checked(s->arena->write_state(0, saved.cur, keep.data(), s->stream), err, err_len, "AMR state restore") ||
      !checked(copy_sync(s->arena->control(), &saved, sizeof(Control), hipMemcpyHostToDevice, s->stream), err, err_len,
               "AMR control restore"))
    return 1;
  if (!ok) return fail(err, err_len, "nonphysical state or nonfinite face flux");
  return 0;
}

}  // extern "C"
