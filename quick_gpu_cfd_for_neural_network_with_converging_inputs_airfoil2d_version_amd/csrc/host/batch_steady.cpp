// batch_steady.cpp -- steady driver over a batch of slots (amr_driver_*, amr_batch_solve).
//
// Every slot runs the check loop of the reference solver executable
// (src/quick_airfoil.cpp:66-95) on its own schedule: its target is its next check time, the
// batch advances until some running slot reaches its target (or fails, or its wall budget
// runs out), then one masked monitor pass checks every slot that is due.  Slots that end are
// reported and go idle (target = their time); the others keep running, so a serving loop can
// refill ended slots between polls.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "amr_hip.h"
#include "steady_check.h"

using cfd::hipamr::SteadyRun;
using Clock = std::chrono::steady_clock;

struct amr_batch_driver {
  amr_batch* batch = nullptr;
  int n = 0;
  std::vector<std::unique_ptr<SteadyRun>> runs;  // null: no solve in progress on the slot
  std::vector<Clock::time_point> started;
  std::vector<char> stepped;                     // the solve has advanced at least once
  std::vector<amr_slot_state> states;
  Clock::time_point reported = Clock::now();  // last progress line
  uint64_t check_rounds = 0;
};

namespace {

int fail(char* err, size_t len, const std::string& message) {
  if (err && len) {
    const size_t k = std::min(message.size(), len - 1);
    std::memcpy(err, message.data(), k);
    err[k] = 0;
  }
  return 1;
}

amr_state_meta meta_of(const amr_slot_state& s) { return amr_state_meta{s.time, s.steps, s.rejected}; }

double seconds_since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

}  // namespace

extern "C" {

int amr_driver_create(amr_batch* b, amr_batch_driver** out, char* err, size_t err_len) {
  if (!out) return fail(err, err_len, "null driver pointer");
  *out = nullptr;
  const int n = amr_batch_size(b);
  if (n < 1) return fail(err, err_len, "invalid AMR batch");
  std::unique_ptr<amr_batch_driver> d(new amr_batch_driver);
  d->batch = b;
  d->n = n;
  d->runs.resize(n);
  d->started.resize(n);
  d->stepped.assign(n, 0);
  d->states.resize(n);
  if (amr_batch_slots(b, d->states.data())) return fail(err, err_len, "invalid AMR batch");
  *out = d.release();
  return 0;
}

int amr_driver_begin(amr_batch_driver* d, int32_t slot, const amr_steady_options* options, char* err,
                     size_t err_len) {
  if (!d || slot < 0 || slot >= d->n || !options || !SteadyRun::valid(*options))
    return fail(err, err_len, "invalid AMR steady options");
  if (amr_batch_slots(d->batch, d->states.data())) return fail(err, err_len, "invalid AMR batch");
  const amr_slot_state& s = d->states[slot];
  if (!s.n_cells) return fail(err, err_len, "steady run on an empty slot");
  if (s.fatal) return fail(err, err_len, "steady run on a failed slot");
  d->runs[slot].reset(new SteadyRun(*options));
  d->started[slot] = Clock::now();
  d->stepped[slot] = 0;
  return 0;
}

int32_t amr_driver_running(const amr_batch_driver* d) {
  if (!d) return 0;
  int32_t k = 0;
  for (const auto& r : d->runs) k += r != nullptr;
  return k;
}

int amr_driver_poll(amr_batch_driver* d, double wall_budget_seconds, amr_steady_result* results, int32_t* done,
                    int32_t* n_done, char* err, size_t err_len) {
  if (!d || !results || !done || !n_done || !std::isfinite(wall_budget_seconds) || wall_budget_seconds < 0)
    return fail(err, err_len, "invalid AMR driver poll");
  const auto t0 = Clock::now();
  const int n = d->n;
  *n_done = 0;
  const auto end_run = [&](int m) {
    results[m] = d->runs[m]->result();
    d->runs[m].reset();
    done[(*n_done)++] = m;
  };
  if (amr_batch_slots(d->batch, d->states.data())) return fail(err, err_len, "invalid AMR batch");
  // Solves that start at or beyond their stop time end without a check.
  for (int m = 0; m < n; ++m)
    if (d->runs[m] && !d->stepped[m] && d->runs[m]->starts_done(meta_of(d->states[m]))) {
      d->runs[m]->stop(meta_of(d->states[m]), seconds_since(d->started[m]));
      end_run(m);
    }
  if (*n_done) return 0;
  std::vector<double> targets(n), frames(6 * n), mach(n), aoa(n), mon(11 * n);
  std::vector<int64_t> maxsteps(n);
  std::vector<int32_t> mask(n), status(n);
  char message[512] = {0};
  for (;;) {
    double budget = INFINITY;
    int running = 0;
    for (int m = 0; m < n; ++m) {
      const amr_slot_state& s = d->states[m];
      if (d->runs[m]) {
        ++running;
        targets[m] = d->runs[m]->target();
        budget = std::min(budget, d->runs[m]->remaining(seconds_since(d->started[m])));
      } else {
        targets[m] = s.time;  // idle: stays finished
      }
      maxsteps[m] = 50000000;
    }
    if (!running) return 0;
    if (wall_budget_seconds > 0) budget = std::min(budget, wall_budget_seconds - seconds_since(t0));
    if (amr_batch_set_targets(d->batch, targets.data(), maxsteps.data(), err, err_len)) return 1;
    const int status_advance =
        amr_batch_advance(d->batch, 2, std::max(.001, budget), d->states.data(), message, sizeof(message));
    if (status_advance == 1) return fail(err, err_len, message);
    for (int m = 0; m < n; ++m) d->stepped[m] |= d->runs[m] != nullptr;
    // Which solves are due: failed ones end, finished or out-of-budget ones are checked.
    int due = 0;
    for (int m = 0; m < n; ++m) {
      mask[m] = 0;
      if (!d->runs[m]) continue;
      const amr_slot_state& s = d->states[m];
      const double el = seconds_since(d->started[m]);
      if (s.fatal || s.rejected) {
        d->runs[m]->fail(meta_of(s), el);
        end_run(m);
      } else if (s.finished || d->runs[m]->remaining(el) <= 0) {
        mask[m] = 1;
        const amr_steady_options& o = d->runs[m]->options();
        for (int i = 0; i < 4; ++i) frames[6 * m + i] = o.region[i];
        frames[6 * m + 4] = o.moment[0];
        frames[6 * m + 5] = o.moment[1];
        mach[m] = o.mach;
        aoa[m] = o.aoa_degrees;
        ++due;
      } else {
        frames[6 * m] = frames[6 * m + 2] = 0;
        frames[6 * m + 1] = frames[6 * m + 3] = -1;  // empty region: no cell counted
        frames[6 * m + 4] = frames[6 * m + 5] = 0;
        mach[m] = 1;
        aoa[m] = 0;
      }
    }
    if (due) {
      if (amr_batch_monitor(d->batch, mask.data(), frames.data(), mach.data(), aoa.data(), mon.data(), status.data(),
                            err, err_len))
        return 1;
      ++d->check_rounds;
      if (seconds_since(d->reported) >= 10) {  // progress of long solves: at most one line per 10 s of checks
        double tmin = INFINITY, tmax = 0;
        for (int m = 0; m < n; ++m)
          if (d->runs[m]) tmin = std::min(tmin, d->states[m].time), tmax = std::max(tmax, d->states[m].time);
        std::fprintf(stderr, "amr-driver: %d solves running, %llu check rounds, t* %.4g..%.4g\n", running,
                     static_cast<unsigned long long>(d->check_rounds), tmin, tmax);
        d->reported = Clock::now();
      }
      for (int m = 0; m < n; ++m) {
        if (!mask[m]) continue;
        const amr_slot_state& s = d->states[m];
        const double el = seconds_since(d->started[m]);
        if (status[m]) {
          d->runs[m]->fail(meta_of(s), el);
          end_run(m);
        } else if (d->runs[m]->check(meta_of(s), mon.data() + 11 * m, el, !s.finished, nullptr)) {
          end_run(m);
        } else if (d->runs[m]->remaining(seconds_since(d->started[m])) <= 0) {
          d->runs[m]->stop(meta_of(s), seconds_since(d->started[m]));  // budget gone at a check
          end_run(m);
        }
      }
    }
    if (*n_done) return 0;
    if (wall_budget_seconds > 0 && seconds_since(t0) >= wall_budget_seconds) return 2;
  }
}

void amr_driver_destroy(amr_batch_driver* d) { delete d; }

int amr_batch_solve(amr_batch* b, const amr_steady_options* options, amr_steady_result* results, char* err,
                    size_t err_len) {
  const int n = amr_batch_size(b);
  if (n < 1 || !options || !results) return fail(err, err_len, "invalid AMR batch solve");
  amr_batch_driver* raw = nullptr;
  if (amr_driver_create(b, &raw, err, err_len)) return 1;
  std::unique_ptr<amr_batch_driver, void (*)(amr_batch_driver*)> d(raw, amr_driver_destroy);
  std::memset(results, 0, n * sizeof(amr_steady_result));
  for (int m = 0; m < n; ++m)
    if (d->states[m].n_cells && amr_driver_begin(d.get(), m, options + m, err, err_len)) return 1;
  std::vector<int32_t> done(n);
  int32_t n_done = 0;
  while (amr_driver_running(d.get()))
    if (amr_driver_poll(d.get(), 0, results, done.data(), &n_done, err, err_len) == 1) return 1;
  return 0;
}

}  // extern "C"
