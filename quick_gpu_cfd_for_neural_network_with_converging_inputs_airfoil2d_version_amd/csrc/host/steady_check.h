// steady_check.h -- the per-solve check logic of the steady / time-accurate driver.
//
// The loop of the reference solver executable (src/quick_airfoil.cpp:66-95): check targets
// 0.01, period, 2 period, ...; at each target the monitor gives the raw residual norms and
// CL/CD/CM_z; the relative coefficient changes are measured against the previous check
// (floors .1 / .01 / .01); `hold` consecutive passing checks after min_time converge.  Any
// rejected step counts as a failure, as in the reference.  One SteadyRun per solve; the
// single-session driver and the batch driver feed it the same way.
#ifndef HIPAMR_STEADY_CHECK_H
#define HIPAMR_STEADY_CHECK_H

#include <algorithm>
#include <cmath>
#include <cstring>

#include "amr_hip.h"

namespace cfd::hipamr {

class SteadyRun {
 public:
  static bool valid(const amr_steady_options& o) {
    return o.stop_time > 0 && std::isfinite(o.stop_time) && o.budget_seconds >= 2 && o.period > 0 &&
           std::isfinite(o.period) && o.hold >= 1 && o.mode >= 0 && o.mode <= 2 && o.mach > 0;
  }

  explicit SteadyRun(const amr_steady_options& o) : o_(o) {
    std::memset(&r_, 0, sizeof(r_));
    r_.reason = AMR_STOP_BUDGET;
    target_ = std::min(.01, o.stop_time);
  }

  const amr_steady_options& options() const { return o_; }
  // Time of the next check.
  double target() const { return target_; }
  // Wall seconds this solve may still spend before its budget expires (one second kept for
  // the final check, as the reference loop does).
  double remaining(double elapsed) const { return o_.budget_seconds - 1 - elapsed; }
  const amr_steady_result& result() const { return r_; }
  int checks() const { return checks_; }

  // A solve that starts at or beyond the stop time performs no check.
  bool starts_done(const amr_state_meta& st) const { return !(st.time < o_.stop_time); }

  // One check at the reached target (or at the state where the budget expired): m is the
  // monitor output (11 values).  Fills the history record when given.  True when the solve ends.
  bool check(const amr_state_meta& st, const double* m, double elapsed, bool budget_expired, amr_steady_check* h) {
    const double cl = m[5], cd = m[6], cm = m[7];
    const double inf = INFINITY;
    const double dcl = previous_ ? std::fabs(cl - prev_[0]) / std::max(.1, std::fabs(cl)) : inf;
    const double dcd = previous_ ? std::fabs(cd - prev_[1]) / std::max(.01, std::fabs(cd)) : inf;
    const double dcm = previous_ ? std::fabs(cm - prev_[2]) / std::max(.01, std::fabs(cm)) : inf;
    const double raw = *std::max_element(m, m + 5);
    if (st.time >= o_.min_time && raw <= o_.raw_tol && dcl <= o_.load_tol && dcd <= o_.load_tol && dcm <= o_.load_tol)
      ++hold_;
    else
      hold_ = 0;
    if (h) {
      h->time = st.time;
      h->steps = st.steps;
      h->rejected = st.rejected;
      h->elapsed = elapsed;
      for (int v = 0; v < 5; ++v) h->residual[v] = m[v];
      h->cl = cl;
      h->cd = cd;
      h->cm = cm;
      h->dcl = dcl;
      h->dcd = dcd;
      h->dcm = dcm;
      h->hold = hold_;
      h->pad = 0;
    }
    for (int v = 0; v < 5; ++v) r_.residual[v] = m[v];
    r_.cl = cl;
    r_.cd = cd;
    r_.cm = cm;
    prev_[0] = cl;
    prev_[1] = cd;
    prev_[2] = cm;
    previous_ = true;
    ++checks_;
    bool done = true;
    if (o_.mode == 2) {
      r_.reason = AMR_STOP_SHORT;
    } else if (o_.mode != 1 && hold_ >= o_.hold) {
      r_.converged = 1;
      r_.reason = AMR_STOP_CONVERGED;
    } else if (budget_expired) {
      r_.reason = AMR_STOP_BUDGET;
    } else if (st.time >= o_.stop_time) {
      r_.reason = AMR_STOP_TIME_CAP;
    } else {
      done = false;
      target_ = checks_ == 1 ? std::min(o_.stop_time, o_.period) : std::min(o_.stop_time, target_ + o_.period);
    }
    finish(st, elapsed);
    return done;
  }

  // The solve failed (fatal device status, a rejected step, or a failed monitor).
  void fail(const amr_state_meta& st, double elapsed) {
    r_.failed = 1;
    r_.reason = AMR_STOP_FAILURE;
    finish(st, elapsed);
  }

  // The solve ended without a check (budget gone before the first target, or start beyond stop).
  void stop(const amr_state_meta& st, double elapsed) {
    r_.reason = AMR_STOP_BUDGET;
    finish(st, elapsed);
  }

 private:
  void finish(const amr_state_meta& st, double elapsed) {
    r_.time = st.time;
    r_.steps = st.steps;
    r_.rejected = st.rejected;
    r_.hold = hold_;
    r_.checks = checks_;
    r_.elapsed_seconds = elapsed;
  }

  amr_steady_options o_;
  amr_steady_result r_;
  double target_ = 0, prev_[3] = {0, 0, 0};
  bool previous_ = false;
  int hold_ = 0, checks_ = 0;
};

}  // namespace cfd::hipamr

#endif
