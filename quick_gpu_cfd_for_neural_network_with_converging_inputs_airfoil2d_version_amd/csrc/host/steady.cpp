// steady.cpp -- steady / time-accurate driver of one device-resident solve (amr_hip_solve).
//
// The check loop of the reference solver executable (src/quick_airfoil.cpp:66-95), with the
// per-check decisions in SteadyRun.  Only amr_hip_advance and amr_hip_monitor touch the
// device: there is no per-check state download.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

#include "amr_hip.h"
#include "steady_check.h"

namespace {
int fail(char* err, size_t len, const std::string& message) {
  if (err && len) {
    const size_t k = std::min(message.size(), len - 1);
    std::memcpy(err, message.data(), k);
    err[k] = 0;
  }
  return 1;
}
}  // namespace

extern "C" int amr_hip_solve(amr_session* s, const amr_steady_options* o, amr_steady_result* r,
                             amr_steady_check* history, int32_t capacity, char* err, size_t err_len) {
  using cfd::hipamr::SteadyRun;
  if (!s || !o || !r || !SteadyRun::valid(*o) || (history && capacity < 0))
    return fail(err, err_len, "invalid AMR steady options");
  const auto begin = std::chrono::steady_clock::now();
  const auto elapsed = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count(); };
  std::memset(r, 0, sizeof(*r));
  amr_state_meta state{};
  if (amr_hip_state(s, &state)) return fail(err, err_len, "invalid AMR session");
  SteadyRun run(*o);
  char message[512] = {0};
  bool ended = false;
  while (run.remaining(elapsed()) > 0 && !run.starts_done(state)) {
    const int status = amr_hip_advance(s, run.target(), 50000000, std::max(.001, run.remaining(elapsed())), &state,
                                       message, sizeof(message));
    double m[11];
    if (status == 1 || state.rejected ||
        amr_hip_monitor(s, o->region, o->moment, o->mach, o->aoa_degrees, m, message, sizeof(message))) {
      run.fail(state, elapsed());
      ended = true;
      break;
    }
    amr_steady_check* h = history && run.checks() < capacity ? history + run.checks() : nullptr;
    if (run.check(state, m, elapsed(), status == 2, h)) {
      ended = true;
      break;
    }
  }
  if (!ended) run.stop(state, elapsed());
  *r = run.result();
  if (r->failed) return fail(err, err_len, message[0] ? message : "AMR steady run failed");
  return 0;
}
