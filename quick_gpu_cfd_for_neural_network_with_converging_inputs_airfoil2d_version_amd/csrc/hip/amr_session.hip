// amr_session.hip -- one device-resident solve behind the C ABI (amr_hip_* in amr_hip.h).
//
// Semantics follow the reference backend (src/static_amr2d_cuda.cu:219-366): create validates,
// uploads and captures one trial as a graph; advance replays it batch_trials times per host
// poll of the 112-byte control record until the target time is reached (0), a fatal code is
// raised (1, last accepted state retained) or the wall budget expires (2); download copies
// the accepted state; stats reports the same counters as AmrGpuStats.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "amr_device.h"
#include "amr_util.h"

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
  hipGraph_t graph_batch = nullptr;      // batch_trials trials in one graph
  hipGraphExec_t exec_batch = nullptr;
  hipEvent_t start = nullptr, end = nullptr;
  Control* host_control = nullptr;  // pinned
  int device = 0;
  uint64_t n_cells = 0, nv = 0;
  double sa_nu_inf = 0;

  ~amr_session() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (exec_batch) (void)hipGraphExecDestroy(exec_batch);
    if (graph_batch) (void)hipGraphDestroy(graph_batch);
    if (start) (void)hipEventDestroy(start);
    if (end) (void)hipEventDestroy(end);
    arena.reset();
    if (host_control) (void)hipHostFree(host_control);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

using Clock = std::chrono::steady_clock;
using hipamr::checked;
using hipamr::copy_sync;
using hipamr::fail;

void to_meta(const Control& c, amr_state_meta* meta) {
  if (!meta) return;
  meta->time = c.time;
  meta->steps = c.steps;
  meta->rejected = c.rejected;
}

int launch_one(amr_session* s) {
  if (s->options.graph && s->exec) return hipGraphLaunch(s->exec, s->stream) == hipSuccess ? 0 : 1;
  s->arena->launch_trial(s->stream);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// `trials` trials: whole poll batches as one graph each (no gap between their trials), the rest
// one by one.
int launch_trials(amr_session* s, int trials) {
  int i = 0;
  if (s->options.graph && s->exec_batch)
    for (; i + s->options.batch_trials <= trials; i += s->options.batch_trials)
      if (hipGraphLaunch(s->exec_batch, s->stream) != hipSuccess) return 1;
  for (; i < trials; ++i)
    if (launch_one(s)) return 1;
  return 0;
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
  s->sa_nu_inf = mesh->transport.sa_nu_inf;
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
  s->stats.pinned_bytes = sizeof(Control) + nv * mesh->n_cells * sizeof(double);
  s->stats.geometry_h2d_bytes = s->arena->geometry_bytes();
  s->stats.state_h2d_bytes = nv * mesh->n_cells * sizeof(double);
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
    // The trials of one poll as a single graph (consecutive graph launches leave a gap between
    // them, consecutive kernels inside one graph none).
    if (options->batch_trials > 1) {
      if (!checked(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), err, err_len, "AMR graph begin"))
        return 1;
      for (int t = 0; t < options->batch_trials; ++t) s->arena->launch_trial(s->stream);
      if (!checked(hipStreamEndCapture(s->stream, &s->graph_batch), err, err_len, "AMR graph end") ||
          !checked(hipGraphInstantiate(&s->exec_batch, s->graph_batch, nullptr, nullptr, 0), err, err_len,
                   "AMR graph instantiate"))
        return 1;
    }
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
  s->arena->launch_prims(s->stream);  // base primitives when the state was written by the host
  for (;;) {
    if (s->options.event_timing && !checked(hipEventRecord(s->start, s->stream), err, err_len, "AMR timing record start"))
      return 1;
    if (launch_trials(s, s->options.batch_trials)) return fail(err, err_len, "AMR trial launch failed");
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
  s->stats.state_d2h_bytes += s->nv * s->n_cells * sizeof(double);
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

int amr_hip_state(const amr_session* s, amr_state_meta* state) {
  if (!s || !state) return 1;
  to_meta(*s->host_control, state);
  return 0;
}

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

int amr_hip_monitor(amr_session* s, const double* region, const double* moment, double mach, double aoa_degrees,
                    double* out, char* err, size_t err_len) {
  if (!s || !region || !moment || !out || !(mach > 0)) return fail(err, err_len, "invalid AMR GPU monitor request");
  // Residual pass on the accepted state with a temporary control, then the reductions.
  Control saved{};
  if (!checked(copy_sync(&saved, s->arena->control(), sizeof(Control), hipMemcpyDeviceToHost, s->stream), err, err_len,
               "AMR control D2H"))
    return 1;
  if (saved.fatal) return fail(err, err_len, "AMR GPU monitor on a failed session");
  Control eval = saved;
  eval.active = 1;
  eval.bad = 0;
  eval.r0_valid = 0;
  eval.dt_min = 0x7ff0000000000000ull;
  Control after{};
  std::vector<double> cells, walls;
  if (!checked(copy_sync(s->arena->control(), &eval, sizeof(Control), hipMemcpyHostToDevice, s->stream), err, err_len,
               "AMR control H2D"))
    return 1;
  const double frame[6] = {region[0], region[1], region[2], region[3], moment[0], moment[1]};
  s->arena->launch_residual(s->stream);
  if (!checked(s->arena->launch_monitor(s->stream, frame), err, err_len, "AMR monitor kernels") ||
      !checked(copy_sync(&after, s->arena->control(), sizeof(Control), hipMemcpyDeviceToHost, s->stream), err, err_len,
               "AMR control D2H") ||
      !checked(s->arena->read_monitor(&cells, &walls, s->stream), err, err_len, "AMR monitor D2H"))
    return 1;
  saved.w_valid = after.bad ? 0 : 1;
  saved.r0_valid = 0;
  saved.dt_min = 0x7ff0000000000000ull;
  if (!checked(copy_sync(s->arena->control(), &saved, sizeof(Control), hipMemcpyHostToDevice, s->stream), err, err_len,
               "AMR control restore"))
    return 1;
  if (after.bad || after.fatal) return fail(err, err_len, "nonphysical state or nonfinite face flux at the monitor");
  double sums[9];
  hipamr::sum_monitor(*s->arena, 0, cells, walls, sums);
  if (!(sums[5] > 0)) return fail(err, err_len, "empty residual region");
  const int nv = static_cast<int>(s->nv);
  for (int v = 0; v < 5; ++v) out[v] = v < nv ? sums[v] / sums[5] : 0.0;
  if (nv == 5) out[4] /= s->sa_nu_inf;
  const double alpha = aoa_degrees * std::acos(-1.) / 180, cs = std::cos(alpha), sn = std::sin(alpha);
  const double qinf = .5 * mach * mach, fx = sums[6], fy = sums[7], mz = sums[8];
  out[5] = (-fx * sn + fy * cs) / qinf;
  out[6] = (fx * cs + fy * sn) / qinf;
  out[7] = mz / qinf;
  out[8] = fx;
  out[9] = fy;
  out[10] = mz;
  return 0;
}

int amr_hip_run(amr_session* s, double tend, int32_t trials, amr_state_meta* state, char* err, size_t err_len) {
  if (!s || !state || trials < 1 || !std::isfinite(tend) || tend < s->host_control->time)
    return fail(err, err_len, "invalid AMR GPU run");
  hipamr::launch_target(s->arena->control(), 1, tend, INT64_MAX, nullptr, nullptr, s->stream);
  s->arena->launch_prims(s->stream);  // base primitives when the state was written by the host
  if (launch_trials(s, trials)) return fail(err, err_len, "AMR trial launch failed");
  if (!checked(hipGetLastError(), err, err_len, "AMR trial kernels") ||
      !checked(copy_sync(s->host_control, s->arena->control(), sizeof(Control), hipMemcpyDeviceToHost, s->stream), err,
               err_len, "AMR control D2H"))
    return 1;
  to_meta(*s->host_control, state);
  if (s->host_control->fatal)
    return fail(err, err_len,
                "AMR GPU stopped with status " + std::to_string(s->host_control->fatal) + "; last accepted state retained");
  return 0;
}

int amr_hip_profile(amr_session* s, double tend, int32_t trials, amr_kernel_time* out, int32_t capacity,
                    int32_t* count, char* err, size_t err_len) {
  if (!s || !out || !count || trials < 1 || capacity < 1 || !std::isfinite(tend) || tend < s->host_control->time)
    return fail(err, err_len, "invalid AMR GPU profile request");
  hipamr::launch_target(s->arena->control(), 1, tend, INT64_MAX, nullptr, nullptr, s->stream);
  s->arena->launch_prims(s->stream);  // base primitives when the state was written by the host
  hipamr::EventHook hook(s->stream);
  for (int i = 0; i < trials; ++i) s->arena->launch_trial(s->stream, &hook);
  if (!checked(hipGetLastError(), err, err_len, "AMR trial kernels") ||
      !checked(copy_sync(s->host_control, s->arena->control(), sizeof(Control), hipMemcpyDeviceToHost, s->stream), err,
               err_len, "AMR control D2H"))
    return 1;
  return hook.collect(out, capacity, count, err, err_len);
}

}  // extern "C"

namespace {
typedef double d2 __attribute__((ext_vector_type(2)));

// Device copy (the STREAM "copy" kernel): each lane moves UNROLL 16-byte vectors per chunk of
// 256 * UNROLL vectors, the block's lanes side by side (coalesced), chunks grid-strided; NT uses
// non-temporal loads and stores (streamed past the caches).
template <int UNROLL, bool NT>
__global__ __launch_bounds__(256) void k_stream_copy(const d2* __restrict__ src, d2* __restrict__ dst, size_t n) {
  const size_t chunk = static_cast<size_t>(256) * UNROLL;
  for (size_t c = static_cast<size_t>(blockIdx.x) * chunk; c < n; c += static_cast<size_t>(gridDim.x) * chunk) {
    d2 v[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const size_t i = c + static_cast<size_t>(u) * 256 + threadIdx.x;
      if (i < n) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const size_t i = c + static_cast<size_t>(u) * 256 + threadIdx.x;
      if (i < n) {
        if (NT) __builtin_nontemporal_store(v[u], dst + i);
        else dst[i] = v[u];
      }
    }
  }
}

// The probe's variants: plain 1 x 16 B per lane, 4 x 16 B, 4 x 16 B non-temporal, 8 x 16 B
// non-temporal.
constexpr int kCopyVariants = 4;
void launch_copy(int variant, const d2* a, d2* b, size_t n, hipStream_t s) {
  const int unroll = variant == 0 ? 1 : (variant == 3 ? 8 : 4);
  const size_t chunks = (n + 256 * unroll - 1) / (256 * unroll);
  const dim3 grid(static_cast<unsigned>(std::min<size_t>(chunks, 256 * 32))), block(256);
  switch (variant) {
    case 0: hipLaunchKernelGGL((k_stream_copy<1, false>), grid, block, 0, s, a, b, n); break;
    case 1: hipLaunchKernelGGL((k_stream_copy<4, false>), grid, block, 0, s, a, b, n); break;
    case 2: hipLaunchKernelGGL((k_stream_copy<4, true>), grid, block, 0, s, a, b, n); break;
    default: hipLaunchKernelGGL((k_stream_copy<8, true>), grid, block, 0, s, a, b, n); break;
  }
}
}  // namespace

extern "C" int amr_hip_copy_variants(uint64_t bytes, int32_t repeats, double* gbps, int32_t capacity, char* err,
                                     size_t err_len) {
  if (!gbps || capacity < kCopyVariants || repeats < 1 || bytes < 4096 || bytes % 16)
    return fail(err, err_len, "invalid copy probe");
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const auto cleanup = [&]() {
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
  };
  if (!checked(hipMalloc(&a, bytes), err, err_len, "copy probe alloc") ||
      !checked(hipMalloc(&b, bytes), err, err_len, "copy probe alloc") ||
      !checked(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), err, err_len, "copy probe stream") ||
      !checked(hipEventCreate(&e0), err, err_len, "copy probe event") ||
      !checked(hipEventCreate(&e1), err, err_len, "copy probe event") ||
      !checked(hipMemsetAsync(a, 0, bytes, s), err, err_len, "copy probe clear")) {
    cleanup();
    return 1;
  }
  const size_t n = bytes / 16;
  const d2* src = static_cast<const d2*>(a);
  d2* dst = static_cast<d2*>(b);
  bool ok = true;
  for (int v = 0; v < kCopyVariants && ok; ++v) {
    launch_copy(v, src, dst, n, s);  // warm-up
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < repeats; ++r) launch_copy(v, src, dst, n, s);
    (void)hipEventRecord(e1, s);
    float ms = 0;
    ok = checked(hipGetLastError(), err, err_len, "copy probe launch") &&
         checked(hipEventSynchronize(e1), err, err_len, "copy probe sync") &&
         checked(hipEventElapsedTime(&ms, e0, e1), err, err_len, "copy probe elapsed");
    if (ok) gbps[v] = 2.0 * static_cast<double>(bytes) / (ms / repeats * 1e-3) / 1e9;
  }
  cleanup();
  return ok ? 0 : 1;
}

extern "C" int amr_hip_copy_bandwidth(uint64_t bytes, int32_t repeats, double* gbps, double* ms_per_copy, char* err,
                                      size_t err_len) {
  if (!gbps || !ms_per_copy) return fail(err, err_len, "invalid copy probe");
  double rates[kCopyVariants] = {};
  if (amr_hip_copy_variants(bytes, repeats, rates, kCopyVariants, err, err_len)) return 1;
  *gbps = *std::max_element(rates, rates + kCopyVariants);
  *ms_per_copy = 2.0 * static_cast<double>(bytes) / (*gbps * 1e9) * 1e3;
  return 0;
}

// ------------------------------------------------------------------------------------------
// amr_hip_run_noalloc: create + advance + download + destroy of one solve, with the arena laid
// out in a file-scope static device array instead of hipMalloc'ed memory.  No device or pinned
// allocation: the caller passes the mesh, parameters and state in host memory; staging goes
// through pageable host memory on the null stream; trials are launched directly (no graph).
namespace {

// 64 MiB per device, held until the process ends: one arena of up to about 70k cells in FP64
// (an LF2 mesh takes 8.5 MiB, a level-7 mesh 26 MiB) or twice that in FP32.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device (its host-side name is not a device address), and
// one lock per device: calls from threads of this process on one device take turns, since each
// rebuilds the arena in the same memory.
std::mutex scratch_lock[kMaxDevices];
void* scratch_ptr[kMaxDevices] = {};

}  // namespace

extern "C" int amr_hip_run_noalloc(const amr_mesh_view* mesh, const amr_params* params, double* u, uint64_t u_len,
                                   amr_state_meta* state, double tend, int64_t maxsteps, double wall_budget_seconds,
                                   int32_t precision, char* err, size_t err_len) {
  if (!mesh || !params || !u || !state) return fail(err, err_len, "invalid AMR GPU configuration");
  MeshInput in{};
  in.view = mesh;
  in.params = *params;
  in.u = u;
  in.meta = *state;
  const uint64_t nv = mesh->transport.sa_model ? 5 : 4;
  std::string message = hipamr::validate_input(in);
  if (message.empty() && u_len != nv * mesh->n_cells) message = "invalid AMR GPU configuration";
  if (message.empty() && precision != 64 && precision != 32) message = "precision must be 64 or 32";
  if (message.empty() && (!std::isfinite(tend) || tend < state->time || maxsteps < 1 ||
                          !std::isfinite(wall_budget_seconds) || wall_budget_seconds < 0))
    message = "invalid AMR GPU advance";
  if (!message.empty()) return fail(err, err_len, message);
  if (amr_hip_device_count() < 1) return fail(err, err_len, "no HIP device available for the AMR backend");
  int device = 0;
  if (!checked(hipGetDevice(&device), err, err_len, "AMR device")) return 1;
  if (device < 0 || device >= kMaxDevices) return fail(err, err_len, "AMR device index beyond the scratch table");
  std::lock_guard<std::mutex> guard(scratch_lock[device]);
  if (!scratch_ptr[device]) {  // once per device, with that device current
    void* p = nullptr;
    if (!checked(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), err, err_len, "AMR scratch address")) return 1;
    scratch_ptr[device] = p;
  }
  const auto t0 = Clock::now();
  const hipStream_t stream = nullptr;
  // The same launch tile as a session of this mesh, so the two give the same bits.
  const int tile = mesh->n_cells < 200000 ? 64 : 256;
  std::unique_ptr<Arena> arena =
      hipamr::make_arena({in}, precision, tile, stream, &message, scratch_ptr[device], kScratchBytes);
  if (!arena) return fail(err, err_len, message);
  Control host{};
  hipamr::launch_target(arena->control(), 1, tend, maxsteps, nullptr, nullptr, stream);
  arena->launch_prims(stream);
  bool expired = false;
  for (;;) {
    for (int trial = 0; trial < 32; ++trial) arena->launch_trial(stream);
    if (!checked(hipGetLastError(), err, err_len, "AMR trial kernels") ||
        !checked(hipMemcpyAsync(&host, arena->control(), sizeof(Control), hipMemcpyDeviceToHost, stream), err, err_len,
                 "AMR control D2H") ||
        !checked(hipStreamSynchronize(stream), err, err_len, "AMR batch sync"))
      return 1;
    if (host.fatal || host.finished) break;
    if (wall_budget_seconds > 0 && std::chrono::duration<double>(Clock::now() - t0).count() >= wall_budget_seconds) {
      expired = true;
      break;
    }
  }
  if (!checked(arena->read_state(0, host.cur, u, stream), err, err_len, "AMR state D2H")) return 1;
  to_meta(host, state);
  if (host.fatal)
    return fail(err, err_len, "AMR GPU stopped with status " + std::to_string(host.fatal) + "; last accepted state retained");
  return expired ? 2 : 0;
}
