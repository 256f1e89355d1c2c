// amr_batch.hip -- many meshes in one device arena behind the C ABI (amr_batch_* in amr_hip.h).
//
// One arena of slots.  With two or more slots the trial runs as independent branches (slot m in
// branch m % branches), each captured as its own graph and replayed on its own stream, so that
// one branch's kernels fill the GPU while another's drain; the streams join at every poll.  The
// host polls the array of control records (112 bytes per slot) every batch_trials trials; in
// amr_batch_advance the polls are pipelined through two pinned snapshots (queue_batch).  Slot
// independence: every kernel reads only its own slot's ranges and control record, the only cross-block
// reductions (minimum dt, monitor partials) are per slot and exact (atomicMin on the bit
// pattern) or done in a fixed block order, so a slot's result does not depend on its
// neighbours.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "amr_device.h"
#include "amr_util.h"

using hipamr::Arena;
using hipamr::checked;
using hipamr::Control;
using hipamr::copy_sync;
using hipamr::fail;
using hipamr::MeshInput;

struct amr_batch {
  std::unique_ptr<Arena> arena;
  amr_batch_options options{};
  amr_gpu_stats stats{};
  hipStream_t stream = nullptr;        // branch 0; every host-side operation is ordered on it
  std::vector<hipStream_t> sides;      // branches 1, 2, ... (branched arenas only)
  hipEvent_t host_done = nullptr;      // recorded on `stream` after host-side work, awaited by `sides`
  bool side_must_wait = false;         // host-side work was queued on `stream` since the last join
  std::vector<hipEvent_t> joined;      // per side stream: its part of a batch is done
  Control* snap[2] = {nullptr, nullptr};  // pinned control snapshots of the pipelined advance
  hipEvent_t snap_done[2] = {nullptr, nullptr};
  std::vector<hipGraph_t> graphs;      // every captured graph (owned)
  std::vector<hipGraphExec_t> execs;   // per branch: one trial
  std::vector<hipGraphExec_t> execs_batch;  // per branch: batch_trials trials in one graph
  // Launch shape while few slots run: at most few_slots running slots replay the trials captured
  // with few_spread threads per cell in the flux launches (the per-trial latency then dominates,
  // and spreading a tile's face items shortens each flux launch); otherwise spread.
  int spread = 1, few_slots = 0, few_spread = 1;
  std::vector<hipGraphExec_t> execs_few, execs_batch_few;
  Control* host_control = nullptr;  // pinned, n slots
  double* host_targets = nullptr;   // pinned, n slots
  long long* host_maxsteps = nullptr;
  int n = 0;

  ~amr_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (hipStream_t s : sides) (void)hipStreamSynchronize(s);
    for (hipGraphExec_t e : execs) (void)hipGraphExecDestroy(e);
    for (hipGraphExec_t e : execs_batch) (void)hipGraphExecDestroy(e);
    for (hipGraphExec_t e : execs_few) (void)hipGraphExecDestroy(e);
    for (hipGraphExec_t e : execs_batch_few) (void)hipGraphExecDestroy(e);
    for (hipGraph_t g : graphs) (void)hipGraphDestroy(g);
    arena.reset();
    if (host_control) (void)hipHostFree(host_control);
    if (host_targets) (void)hipHostFree(host_targets);
    if (host_maxsteps) (void)hipHostFree(host_maxsteps);
    if (host_done) (void)hipEventDestroy(host_done);
    for (hipEvent_t e : joined) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; ++i) {
      if (snap_done[i]) (void)hipEventDestroy(snap_done[i]);
      if (snap[i]) (void)hipHostFree(snap[i]);
    }
    for (hipStream_t s : sides) (void)hipStreamDestroy(s);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

using Clock = std::chrono::steady_clock;

bool options_ok(const amr_batch_options* o) {
  return o && (o->precision == 32 || o->precision == 64) && (o->tile == 64 || o->tile == 128 || o->tile == 256) &&
         o->batch_trials >= 1 && o->batch_trials <= 1024 && (o->graph == 0 || o->graph == 1);
}

hipamr::SlotCapacity to_cap(const amr_slot_capacity& c) {
  return hipamr::SlotCapacity{c.cells, c.faces, c.walls, c.csr, c.terms, c.groups, c.members};
}

// Whether the next trials run with the few-slot launch shape (`running` slots are running).
bool few_shape(const amr_batch* b, int running) {
  return b->few_slots > 0 && b->few_spread != b->spread && running <= b->few_slots;
}

// Join the branches, then read every control record.
int poll(amr_batch* b, char* err, size_t err_len) {
  for (hipStream_t s : b->sides)
    if (!checked(hipStreamSynchronize(s), err, err_len, "AMR branch join")) return 1;
  if (!checked(copy_sync(b->host_control, b->arena->control(), b->n * sizeof(Control), hipMemcpyDeviceToHost,
                         b->stream),
               err, err_len, "AMR control D2H"))
    return 1;
  b->side_must_wait = false;
  b->stats.control_d2h_bytes += b->n * sizeof(Control);
  ++b->stats.status_polls;
  return 0;
}

void fill_states(const amr_batch* b, amr_slot_state* out) {
  if (!out) return;
  const auto& slots = b->arena->slots();
  for (int m = 0; m < b->n; ++m) {
    const Control& c = b->host_control[m];
    amr_slot_state& s = out[m];
    s.time = c.time;
    s.target = c.target;
    s.steps = c.steps;
    s.rejected = c.rejected;
    s.n_cells = slots[m].n_cells;
    s.finished = c.finished;
    s.fatal = c.fatal;
    s.pad = 0;
  }
}

hipStream_t branch_stream(const amr_batch* b, int j) { return j == 0 ? b->stream : b->sides[j - 1]; }

// One batch of the pipelined advance: batch_trials trials of every branch with a running slot
// in the host's latest view (a slot starts running only through the host, between advances,
// so a branch without one stays idle), then, on the main stream once every branch is done, a
// copy of all control records into snapshot `k` (snap_done[k] completes with it).  The side
// streams start the batch only after the work already queued on the main stream (the previous
// snapshot's copy, host-side writes), so a slot's record is never copied while one of its
// trials runs.
int queue_batch(amr_batch* b, int k) {
  if (!b->sides.empty()) {
    if (hipEventRecord(b->host_done, b->stream) != hipSuccess) return 1;
    for (hipStream_t s : b->sides)
      if (hipStreamWaitEvent(s, b->host_done, 0) != hipSuccess) return 1;
    b->side_must_wait = false;
  }
  const int branches = b->arena->n_branches();
  std::vector<char> busy(branches, branches == 1);
  int running = 0;
  for (int m = 0; m < b->n; ++m) {
    const Control& c = b->host_control[m];
    const bool on = !c.finished && !c.fatal;
    running += on;
    if (branches > 1) busy[m % branches] |= on;
  }
  const bool few = few_shape(b, running);
  const std::vector<hipGraphExec_t>& poll_graphs = few ? b->execs_batch_few : b->execs_batch;
  if (poll_graphs.empty()) b->arena->set_flux_spread(few ? b->few_spread : b->spread);
  for (int j = 0; j < branches; ++j) {
    if (!busy[j]) continue;
    hipStream_t s = branch_stream(b, j);
    if (!poll_graphs.empty()) {  // the batch's trials as one graph
      if (hipGraphLaunch(poll_graphs[j], s) != hipSuccess) return 1;
    } else {
      for (int trial = 0; trial < b->options.batch_trials; ++trial) b->arena->launch_trial_branch(j, s);
    }
  }
  for (size_t j = 0; j < b->sides.size(); ++j)
    if (hipEventRecord(b->joined[j], b->sides[j]) != hipSuccess ||
        hipStreamWaitEvent(b->stream, b->joined[j], 0) != hipSuccess)
      return 1;
  if (hipMemcpyAsync(b->snap[k], b->arena->control(), b->n * sizeof(Control), hipMemcpyDeviceToHost, b->stream) !=
          hipSuccess ||
      hipEventRecord(b->snap_done[k], b->stream) != hipSuccess)
    return 1;
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Waits for snapshot k and makes it the host's view of the control records.
int take_snapshot(amr_batch* b, int k, char* err, size_t err_len) {
  if (!checked(hipEventSynchronize(b->snap_done[k]), err, err_len, "AMR control snapshot")) return 1;
  std::copy(b->snap[k], b->snap[k] + b->n, b->host_control);
  b->stats.control_d2h_bytes += b->n * sizeof(Control);
  ++b->stats.status_polls;
  return 0;
}

// One trial of every slot: each branch on its own stream.  The side streams first wait for
// host-side work still queued on the main stream (a monitor's control restore).
int launch_one(amr_batch* b) {
  if (!b->sides.empty() && b->side_must_wait) {
    if (hipEventRecord(b->host_done, b->stream) != hipSuccess) return 1;
    for (hipStream_t s : b->sides)
      if (hipStreamWaitEvent(s, b->host_done, 0) != hipSuccess) return 1;
    b->side_must_wait = false;
  }
  const int branches = b->arena->n_branches();
  int running = 0;
  for (int m = 0; m < b->n; ++m) running += !b->host_control[m].finished && !b->host_control[m].fatal;
  const bool few = few_shape(b, running);
  const std::vector<hipGraphExec_t>& graphs = few ? b->execs_few : b->execs;
  if (graphs.empty()) b->arena->set_flux_spread(few ? b->few_spread : b->spread);
  for (int j = 0; j < branches; ++j) {
    hipStream_t s = branch_stream(b, j);
    if (!graphs.empty()) {
      if (hipGraphLaunch(graphs[j], s) != hipSuccess) return 1;
    } else {
      b->arena->launch_trial_branch(j, s);
    }
  }
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// The trial graphs of every branch at the arena's current launch shape: one trial (into
// `one`), and the trials of one poll as one graph (into `poll`).
int capture_graphs(amr_batch* b, std::vector<hipGraphExec_t>* one, std::vector<hipGraphExec_t>* poll, char* err,
                   size_t err_len) {
  const int branches = b->arena->n_branches();
  b->stats.graph_nodes_per_trial = 0;
  for (int j = 0; j < branches; ++j) {
    hipStream_t s = branch_stream(b, j);
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    if (!checked(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), err, err_len, "AMR graph begin"))
      return 1;
    b->arena->launch_trial_branch(j, s);
    if (!checked(hipStreamEndCapture(s, &g), err, err_len, "AMR graph end")) return 1;
    b->graphs.push_back(g);
    if (!checked(hipGraphInstantiate(&e, g, nullptr, nullptr, 0), err, err_len, "AMR graph instantiate")) return 1;
    one->push_back(e);
    size_t nodes = 0;
    if (!checked(hipGraphGetNodes(g, nullptr, &nodes), err, err_len, "AMR graph node count")) return 1;
    b->stats.graph_nodes_per_trial += nodes;
  }
  // The trials of one poll as a single graph per branch: consecutive graph launches on a
  // stream leave a gap between them (about 6 us on one LF2 mesh, where a trial takes 70 us),
  // consecutive kernels inside one graph none.
  if (b->options.batch_trials > 1)
    for (int j = 0; j < branches; ++j) {
      hipStream_t s = branch_stream(b, j);
      hipGraph_t g = nullptr;
      hipGraphExec_t e = nullptr;
      if (!checked(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), err, err_len, "AMR graph begin"))
        return 1;
      for (int trial = 0; trial < b->options.batch_trials; ++trial) b->arena->launch_trial_branch(j, s);
      if (!checked(hipStreamEndCapture(s, &g), err, err_len, "AMR graph end")) return 1;
      b->graphs.push_back(g);
      if (!checked(hipGraphInstantiate(&e, g, nullptr, nullptr, 0), err, err_len, "AMR graph instantiate"))
        return 1;
      poll->push_back(e);
    }
  return 0;
}

// Drops the captured graphs (the streams are idle).
void release_graphs(amr_batch* b) {
  for (auto* v : {&b->execs, &b->execs_batch, &b->execs_few, &b->execs_batch_few}) {
    for (hipGraphExec_t e : *v) (void)hipGraphExecDestroy(e);
    v->clear();
  }
  for (hipGraph_t g : b->graphs) (void)hipGraphDestroy(g);
  b->graphs.clear();
}

// The graphs of both launch shapes (the few-slot set only when it differs).  On failure no
// graph is kept (trials are then launched directly), so no partial set is ever replayed.
int capture_all(amr_batch* b, char* err, size_t err_len) {
  bool good = b->arena->set_flux_spread(b->spread);
  if (!good) fail(err, err_len, "invalid AMR launch shape");
  good = good && !capture_graphs(b, &b->execs, &b->execs_batch, err, err_len);
  if (good && b->few_slots > 0 && b->few_spread != b->spread) {
    const size_t nodes = b->stats.graph_nodes_per_trial;
    good = b->arena->set_flux_spread(b->few_spread);
    if (!good) fail(err, err_len, "invalid AMR launch shape");
    good = good && !capture_graphs(b, &b->execs_few, &b->execs_batch_few, err, err_len);
    b->stats.graph_nodes_per_trial = nodes;
  }
  b->arena->set_flux_spread(b->spread);
  if (!good) release_graphs(b);
  return good ? 0 : 1;
}


// Stream, pinned control mirrors, optional graph; the arena is already built.
int finish_setup(amr_batch* b, Clock::time_point t0, char* err, size_t err_len) {
  const int n = b->n;
  if (!checked(hipHostMalloc(reinterpret_cast<void**>(&b->host_control), n * sizeof(Control), hipHostMallocDefault),
               err, err_len, "AMR pinned control") ||
      !checked(hipHostMalloc(reinterpret_cast<void**>(&b->host_targets), n * sizeof(double), hipHostMallocDefault), err,
               err_len, "AMR pinned targets") ||
      !checked(hipHostMalloc(reinterpret_cast<void**>(&b->host_maxsteps), n * sizeof(long long), hipHostMallocDefault),
               err, err_len, "AMR pinned step caps"))
    return 1;
  if (poll(b, err, err_len)) return 1;
  b->stats.allocated_bytes = b->arena->device_bytes();
  b->stats.geometry_h2d_bytes = b->arena->geometry_bytes();
  b->stats.pinned_bytes = n * (3 * sizeof(Control) + sizeof(double) + sizeof(long long));
  const int branches = b->arena->n_branches();
  for (int j = 1; j < branches; ++j) {
    hipStream_t s = nullptr;
    hipEvent_t e = nullptr;
    if (!checked(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), err, err_len, "AMR branch stream")) return 1;
    b->sides.push_back(s);
    if (!checked(hipEventCreateWithFlags(&e, hipEventDisableTiming), err, err_len, "AMR branch event")) return 1;
    b->joined.push_back(e);
  }
  for (int i = 0; i < 2; ++i)
    if (!checked(hipHostMalloc(reinterpret_cast<void**>(&b->snap[i]), n * sizeof(Control), hipHostMallocDefault), err,
                 err_len, "AMR pinned control snapshot") ||
        !checked(hipEventCreateWithFlags(&b->snap_done[i], hipEventDisableTiming), err, err_len, "AMR snapshot event"))
      return 1;
  if (branches > 1 &&
      !checked(hipEventCreateWithFlags(&b->host_done, hipEventDisableTiming), err, err_len, "AMR branch event"))
    return 1;
  if (b->options.graph && capture_all(b, err, err_len)) return 1;
  b->stats.setup_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
  return 0;
}

bool slot_ok(const amr_batch* b, int32_t slot) { return b && slot >= 0 && slot < b->n; }

}  // namespace

extern "C" {

int amr_batch_capacity(const amr_mesh_view* mesh, int32_t tile, amr_slot_capacity* capacity,
                       amr_batch_features* features) {
  if (!mesh || !capacity || (tile != 64 && tile != 128 && tile != 256) || mesh->n_cells == 0 ||
      mesh->n_cells > (1u << 29) || !mesh->face_offsets || !mesh->group_offsets || !mesh->member_offsets)
    return 1;
  const hipamr::SlotCapacity c = hipamr::capacity_of(*mesh, tile);
  *capacity = amr_slot_capacity{c.cells, c.faces, c.walls, c.csr, c.terms, c.groups, c.members, 0};
  if (features) {
    const hipamr::ArenaFeatures f = hipamr::features_of(*mesh);
    *features = amr_batch_features{f.laminar, f.sa, f.small, f.visc_group};
  }
  return 0;
}

uint64_t amr_batch_bytes(const amr_slot_capacity* capacities, int32_t n, const amr_batch_features* features,
                         int32_t precision, int32_t tile) {
  if (!capacities || n < 1 || !features || (precision != 32 && precision != 64) ||
      (tile != 64 && tile != 128 && tile != 256))
    return 0;
  std::vector<hipamr::SlotCapacity> caps;
  for (int32_t m = 0; m < n; ++m) caps.push_back(to_cap(capacities[m]));
  const hipamr::ArenaFeatures f{features->laminar, features->sa, features->small, features->visc_group};
  return hipamr::arena_bytes(caps, f, precision, tile);
}

int amr_batch_create(const amr_mesh_view* meshes, const amr_params* params, const double* const* u,
                     const amr_state_meta* initial, int32_t n, const amr_batch_options* options, amr_batch** out,
                     char* err, size_t err_len) {
  if (!out) return fail(err, err_len, "null batch pointer");
  *out = nullptr;
  if (!meshes || !params || !u || n < 1 || !options_ok(options)) return fail(err, err_len, "invalid AMR batch");
  if (amr_hip_device_count() < 1) return fail(err, err_len, "no HIP device available for the AMR backend");
  const auto t0 = Clock::now();
  std::vector<MeshInput> inputs(n);
  uint64_t state_bytes = 0;
  for (int32_t m = 0; m < n; ++m) {
    inputs[m].view = meshes + m;
    inputs[m].params = params[m];
    inputs[m].u = u[m];
    inputs[m].meta = initial ? initial[m] : amr_state_meta{0, 0, 0};
    const std::string e = hipamr::validate_input(inputs[m]);
    if (!e.empty()) return fail(err, err_len, "mesh " + std::to_string(m) + ": " + e);
    if ((meshes[m].transport.sa_model != 0) != (meshes[0].transport.sa_model != 0))
      return fail(err, err_len, "every mesh of a batch has the same number of conserved variables");
    state_bytes += (meshes[m].transport.sa_model ? 5 : 4) * meshes[m].n_cells * sizeof(double);
  }
  std::unique_ptr<amr_batch> b(new amr_batch);
  b->options = *options;
  b->n = n;
  if (!checked(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), err, err_len, "AMR stream")) return 1;
  std::string message;
  b->arena = hipamr::make_arena(inputs, options->precision, options->tile, b->stream, &message);
  if (!b->arena) return fail(err, err_len, message);
  b->stats.state_h2d_bytes = state_bytes;
  b->stats.control_h2d_bytes = n * sizeof(Control);
  if (finish_setup(b.get(), t0, err, err_len)) return 1;
  *out = b.release();
  return 0;
}

int amr_batch_create_slots(const amr_slot_capacity* capacities, int32_t n, const amr_batch_features* features,
                           const amr_batch_options* options, amr_batch** out, char* err, size_t err_len) {
  if (!out) return fail(err, err_len, "null batch pointer");
  *out = nullptr;
  if (!capacities || !features || n < 1 || !options_ok(options)) return fail(err, err_len, "invalid AMR batch");
  if (amr_hip_device_count() < 1) return fail(err, err_len, "no HIP device available for the AMR backend");
  const auto t0 = Clock::now();
  std::vector<hipamr::SlotCapacity> caps;
  for (int32_t m = 0; m < n; ++m) caps.push_back(to_cap(capacities[m]));
  const hipamr::ArenaFeatures f{features->laminar != 0, features->sa != 0, features->small != 0,
                                features->visc_group != 0};
  std::unique_ptr<amr_batch> b(new amr_batch);
  b->options = *options;
  b->n = n;
  if (!checked(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), err, err_len, "AMR stream")) return 1;
  std::string message;
  b->arena = hipamr::make_slot_arena(caps, f, options->precision, options->tile, b->stream, &message);
  if (!b->arena) return fail(err, err_len, message);
  b->stats.control_h2d_bytes = n * sizeof(Control);
  if (finish_setup(b.get(), t0, err, err_len)) return 1;
  *out = b.release();
  return 0;
}

int amr_batch_admit(amr_batch* b, int32_t slot, const amr_mesh_view* mesh, const amr_params* params,
                    const double* u, uint64_t u_len, const amr_state_meta* initial, char* err, size_t err_len) {
  if (!slot_ok(b, slot) || !mesh || !params || !u) return fail(err, err_len, "invalid AMR batch admission");
  const uint64_t nv = mesh->transport.sa_model ? 5 : 4;
  if (u_len != nv * mesh->n_cells) return fail(err, err_len, "state length does not match the mesh");
  MeshInput in{};
  in.view = mesh;
  in.params = *params;
  in.u = u;
  in.meta = initial ? *initial : amr_state_meta{0, 0, 0};
  const std::string e = b->arena->admit(slot, in, b->stream);
  if (!e.empty()) {
    poll(b, nullptr, 0);
    return fail(err, err_len, e);
  }
  b->stats.state_h2d_bytes += u_len * sizeof(double);
  b->stats.control_h2d_bytes += sizeof(Control);
  return poll(b, err, err_len);
}

int amr_batch_release(amr_batch* b, int32_t slot, char* err, size_t err_len) {
  if (!slot_ok(b, slot)) return fail(err, err_len, "invalid AMR batch slot");
  const std::string e = b->arena->release(slot, b->stream);
  if (!e.empty()) return fail(err, err_len, e);
  return poll(b, err, err_len);
}

int amr_batch_set_targets(amr_batch* b, const double* targets, const int64_t* maxsteps, char* err, size_t err_len) {
  if (!b || !targets) return fail(err, err_len, "invalid AMR batch targets");
  const auto& slots = b->arena->slots();
  for (int m = 0; m < b->n; ++m) {
    if (std::isnan(targets[m]) || (maxsteps && maxsteps[m] < 1)) return fail(err, err_len, "invalid AMR batch targets");
    // An empty slot stays finished whatever its entry holds.
    b->host_targets[m] = slots[m].n_cells ? targets[m] : -INFINITY;
    b->host_maxsteps[m] = maxsteps ? maxsteps[m] : INT64_MAX;
  }
  if (!checked(hipMemcpyAsync(b->arena->target_array(), b->host_targets, b->n * sizeof(double),
                              hipMemcpyHostToDevice, b->stream),
               err, err_len, "AMR targets H2D") ||
      !checked(hipMemcpyAsync(b->arena->maxsteps_array(), b->host_maxsteps, b->n * sizeof(long long),
                              hipMemcpyHostToDevice, b->stream),
               err, err_len, "AMR step caps H2D"))
    return 1;
  hipamr::launch_target(b->arena->control(), b->n, 0.0, 0, b->arena->target_array(), b->arena->maxsteps_array(),
                        b->stream);
  b->arena->launch_prims(b->stream);  // base primitives of slots whose state the host wrote
  if (!checked(hipGetLastError(), err, err_len, "AMR target kernel")) return 1;
  return poll(b, err, err_len);
}

int amr_batch_advance(amr_batch* b, int32_t until, double wall_budget_seconds, amr_slot_state* states, char* err,
                      size_t err_len) {
  if (!b || until < 0 || until > 2 || !std::isfinite(wall_budget_seconds) || wall_budget_seconds < 0)
    return fail(err, err_len, "invalid AMR batch advance");
  const auto t0 = Clock::now();
  auto last = t0;
  std::vector<char> running(b->n);
  int n_running = 0;
  for (int m = 0; m < b->n; ++m) {
    const Control& c = b->host_control[m];
    running[m] = !c.finished && !c.fatal;
    n_running += running[m];
  }
  bool expired = false;
  // until = 2: once a slot has ended, keep going until an eighth of the running slots have
  // ended or a few more polls have passed, so that one check serves several slots.  The wait
  // shrinks with the number of running slots (8 polls at 256 and more, none at 32 and fewer):
  // a waiting slot is a larger share of the work when few are running.
  const int group = std::max(1, n_running / 8);
  const int wait_polls = std::min(8, n_running / 32);
  int polls = 0, first_end = -1;
  // Pipelined: batch k + 1 is queued before the host waits for batch k's snapshot, so the GPU
  // never waits for the host; on exit the batch in flight completes and its snapshot is the
  // state returned.  A slot stops at its own target, so the extra batch changes no trajectory.
  int k = 0;
  if (n_running && queue_batch(b, k)) return fail(err, err_len, "AMR trial launch failed");
  while (n_running) {
    if (queue_batch(b, k ^ 1)) return fail(err, err_len, "AMR trial launch failed");
    if (take_snapshot(b, k, err, err_len)) return 1;
    k ^= 1;
    int still = 0;
    for (int m = 0; m < b->n; ++m) {
      const Control& c = b->host_control[m];
      still += running[m] && !c.finished && !c.fatal;
    }
    const auto now = Clock::now();
    if (std::chrono::duration<double>(now - last).count() >= 2) {
      std::fprintf(stderr, "amr-batch: %d of %d slots running, polls=%llu\n", still, n_running,
                   static_cast<unsigned long long>(b->stats.status_polls));
      last = now;
    }
    ++polls;
    const int ended = n_running - still;
    if (still == 0 || (until == 1 && ended > 0)) break;
    if (until == 2 && ended > 0) {
      if (first_end < 0) first_end = polls;
      if (ended >= group || polls - first_end >= wait_polls) break;
    }
    if (wall_budget_seconds > 0 && std::chrono::duration<double>(now - t0).count() >= wall_budget_seconds) {
      expired = true;
      break;
    }
  }
  if (polls && take_snapshot(b, k, err, err_len)) return 1;  // the batch still in flight
  b->stats.advance_seconds += std::chrono::duration<double>(Clock::now() - t0).count();
  fill_states(b, states);
  return expired ? 2 : 0;
}

int amr_batch_run(amr_batch* b, int32_t trials, amr_slot_state* states, char* err, size_t err_len) {
  if (!b || trials < 1) return fail(err, err_len, "invalid AMR batch run");
  int i = 0;
  int running = 0;
  for (int m = 0; m < b->n; ++m) running += !b->host_control[m].finished && !b->host_control[m].fatal;
  const std::vector<hipGraphExec_t>& poll_graphs = few_shape(b, running) ? b->execs_batch_few : b->execs_batch;
  if (!poll_graphs.empty() && trials >= b->options.batch_trials) {
    if (!b->sides.empty() && b->side_must_wait) {  // as launch_one: side streams after host-side work
      if (hipEventRecord(b->host_done, b->stream) != hipSuccess) return fail(err, err_len, "AMR trial launch failed");
      for (hipStream_t s : b->sides)
        if (hipStreamWaitEvent(s, b->host_done, 0) != hipSuccess) return fail(err, err_len, "AMR trial launch failed");
      b->side_must_wait = false;
    }
    for (; i + b->options.batch_trials <= trials; i += b->options.batch_trials)
      for (int j = 0; j < b->arena->n_branches(); ++j)
        if (hipGraphLaunch(poll_graphs[j], branch_stream(b, j)) != hipSuccess)
          return fail(err, err_len, "AMR trial launch failed");
  }
  for (; i < trials; ++i)
    if (launch_one(b)) return fail(err, err_len, "AMR trial launch failed");
  if (!checked(hipGetLastError(), err, err_len, "AMR trial kernels") || poll(b, err, err_len)) return 1;
  fill_states(b, states);
  return 0;
}

int amr_batch_set_launch_shape(amr_batch* b, int32_t flux_spread, int32_t few_slots, int32_t few_flux_spread, char* err,
                               size_t err_len) {
  const auto ok_spread = [&](int s) { return s == 1 || s == 2; };
  if (!b || !ok_spread(flux_spread) || !ok_spread(few_flux_spread) || few_slots < 0 ||
      b->options.tile * std::max(flux_spread, few_flux_spread) > 256)
    return fail(err, err_len, "invalid AMR launch shape (threads per cell 1 or 2, at most 256 per block)");
  if (!checked(hipStreamSynchronize(b->stream), err, err_len, "AMR launch shape")) return 1;
  for (hipStream_t s : b->sides)
    if (!checked(hipStreamSynchronize(s), err, err_len, "AMR launch shape")) return 1;
  release_graphs(b);
  b->spread = flux_spread;
  b->few_slots = few_slots;
  b->few_spread = few_flux_spread;
  if (b->options.graph) return capture_all(b, err, err_len);
  b->arena->set_flux_spread(flux_spread);
  return 0;
}

int amr_batch_slots(const amr_batch* b, amr_slot_state* states) {
  if (!b || !states) return 1;
  fill_states(b, states);
  return 0;
}

int32_t amr_batch_size(const amr_batch* b) { return b ? b->n : 0; }

int amr_batch_download(amr_batch* b, int32_t slot, double* u, uint64_t u_len, amr_state_meta* state, char* err,
                       size_t err_len) {
  if (!slot_ok(b, slot) || !u) return fail(err, err_len, "invalid AMR batch download");
  const hipamr::MeshSlot& s = b->arena->slots()[slot];
  if (s.n_cells == 0 || u_len != static_cast<uint64_t>(s.nv) * s.n_cells)
    return fail(err, err_len, "invalid AMR batch download");
  if (poll(b, err, err_len)) return 1;
  const Control& c = b->host_control[slot];
  if (!checked(b->arena->read_state(slot, c.cur, u, b->stream), err, err_len, "AMR state D2H")) return 1;
  if (state) *state = amr_state_meta{c.time, c.steps, c.rejected};
  b->stats.state_d2h_bytes += u_len * sizeof(double);
  ++b->stats.full_downloads;
  return 0;
}

int amr_batch_monitor(amr_batch* b, const int32_t* mask, const double* frames, const double* mach,
                      const double* aoa_degrees, double* out, int32_t* status, char* err, size_t err_len) {
  if (!b || !mask || !frames || !mach || !aoa_degrees || !out || !status)
    return fail(err, err_len, "invalid AMR batch monitor request");
  const int n = b->n;
  const auto& slots = b->arena->slots();
  if (poll(b, err, err_len)) return 1;
  std::vector<Control> saved(b->host_control, b->host_control + n), eval = saved;
  std::vector<char> use(n, 0);
  int any = 0;
  for (int m = 0; m < n; ++m) {
    use[m] = mask[m] && slots[m].n_cells && !saved[m].fatal && mach[m] > 0;
    any += use[m];
    Control& e = eval[m];
    e.active = use[m];
    if (use[m]) {
      e.bad = 0;
      e.r0_valid = 0;
      e.dt_min = hipamr::kInfBitsHost;
    }
  }
  for (int m = 0; m < n; ++m) {
    status[m] = 3;
    for (int i = 0; i < 11; ++i) out[11 * m + i] = NAN;
  }
  if (!any) return 0;
  std::vector<double> cells, walls;
  // One synchronisation for the whole evaluation: control in, residual, monitor sums and
  // control out are ordered on the stream; read_monitor waits for all of them.
  std::copy(eval.begin(), eval.end(), b->host_control);
  if (!checked(hipMemcpyAsync(b->arena->control(), b->host_control, n * sizeof(Control), hipMemcpyHostToDevice,
                              b->stream),
               err, err_len, "AMR control H2D"))
    return 1;
  b->arena->launch_residual(b->stream);
  if (!checked(b->arena->launch_monitor(b->stream, frames), err, err_len, "AMR monitor kernels") ||
      !checked(hipMemcpyAsync(b->host_control, b->arena->control(), n * sizeof(Control), hipMemcpyDeviceToHost,
                              b->stream),
               err, err_len, "AMR control D2H") ||
      !checked(b->arena->read_monitor(&cells, &walls, b->stream), err, err_len, "AMR monitor D2H"))
    return 1;
  std::vector<Control> restore = saved;
  for (int m = 0; m < n; ++m) {
    if (!use[m]) continue;
    const bool bad = b->host_control[m].bad || b->host_control[m].fatal;
    restore[m].w_valid = bad ? 0 : 1;
    restore[m].r0_valid = 0;
    restore[m].dt_min = hipamr::kInfBitsHost;
    if (bad) {
      status[m] = 1;
      continue;
    }
    double sums[9];
    hipamr::sum_monitor(*b->arena, m, cells, walls, sums);
    if (!(sums[5] > 0)) {
      status[m] = 2;
      continue;
    }
    double* o = out + 11 * m;
    const int nv = slots[m].nv;
    for (int v = 0; v < 5; ++v) o[v] = v < nv ? sums[v] / sums[5] : 0.0;
    if (nv == 5) o[4] /= b->arena->sa_nu_inf(m);
    const double alpha = aoa_degrees[m] * std::acos(-1.) / 180, cs = std::cos(alpha), sn = std::sin(alpha);
    const double qinf = .5 * mach[m] * mach[m], fx = sums[6], fy = sums[7], mz = sums[8];
    o[5] = (-fx * sn + fy * cs) / qinf;
    o[6] = (fx * cs + fy * sn) / qinf;
    o[7] = mz / qinf;
    o[8] = fx;
    o[9] = fy;
    o[10] = mz;
    status[m] = 0;
  }
  // The restore is ordered before every later use of the control records on the stream; the
  // pinned host copy is only overwritten by later device-to-host copies on the same stream.
  std::copy(restore.begin(), restore.end(), b->host_control);
  if (!checked(hipMemcpyAsync(b->arena->control(), b->host_control, n * sizeof(Control), hipMemcpyHostToDevice,
                              b->stream),
               err, err_len, "AMR control restore"))
    return 1;
  b->side_must_wait = true;
  return 0;
}

int amr_batch_profile(amr_batch* b, int32_t trials, amr_kernel_time* out, int32_t capacity, int32_t* count,
                      char* err, size_t err_len) {
  if (!b || !out || !count || trials < 1 || capacity < 1) return fail(err, err_len, "invalid AMR batch profile");
  hipamr::EventHook hook(b->stream);
  for (int i = 0; i < trials; ++i) b->arena->launch_trial(b->stream, &hook);
  if (!checked(hipGetLastError(), err, err_len, "AMR trial kernels") || poll(b, err, err_len)) return 1;
  return hook.collect(out, capacity, count, err, err_len);
}

void amr_batch_stats(const amr_batch* b, amr_gpu_stats* out) {
  if (!out) return;
  if (!b) {
    std::memset(out, 0, sizeof(*out));
    return;
  }
  *out = b->stats;
}

void amr_batch_destroy(amr_batch* b) { delete b; }

}  // extern "C"
