// amr_util.h -- error reporting, ordered copies and per-kernel event timing shared by the
// single-mesh session and the batch entry points.
#ifndef HIPAMR_AMR_UTIL_H
#define HIPAMR_AMR_UTIL_H

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "amr_arena.h"
#include "amr_hip.h"

namespace hipamr {

// Copy the message into err (truncated, NUL-terminated) and return status 1.
inline int fail(char* err, size_t len, const std::string& message) {
  if (err && len) {
    const size_t k = message.size() < len - 1 ? message.size() : len - 1;
    std::memcpy(err, message.data(), k);
    err[k] = 0;
  }
  return 1;
}

inline bool checked(hipError_t status, char* err, size_t len, const char* operation) {
  if (status == hipSuccess) return true;
  fail(err, len, std::string(operation) + ": " + hipGetErrorString(status));
  return false;
}

// Synchronous copy ordered on `stream` (a non-blocking stream does not order with the null stream).
inline hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t stream) {
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, stream);
  return e != hipSuccess ? e : hipStreamSynchronize(stream);
}

constexpr unsigned long long kInfBitsHost = 0x7ff0000000000000ull;

// Records a hipEvent pair around every kernel launch on a stream.
struct EventHook final : LaunchHook {
  hipStream_t stream;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> marks;
  hipError_t status = hipSuccess;
  explicit EventHook(hipStream_t s) : stream(s) {}
  ~EventHook() override {
    for (auto& m : marks) {
      if (m.second.first) (void)hipEventDestroy(m.second.first);
      if (m.second.second) (void)hipEventDestroy(m.second.second);
    }
  }
  void before(const char* name) override {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) status = hipErrorOutOfMemory;
    marks.push_back({name, {a, b}});
    if (a) (void)hipEventRecord(a, stream);
  }
  void after(const char*) override {
    if (marks.back().second.second) (void)hipEventRecord(marks.back().second.second, stream);
  }

  // Aggregate by kernel name in first-launch order (the stream must have been synchronised).
  int collect(amr_kernel_time* out, int32_t capacity, int32_t* count, char* err, size_t err_len) const {
    if (!checked(status, err, err_len, "AMR profile events")) return 1;
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, int64_t>> totals;
    for (const auto& m : marks) {
      float ms = 0;
      if (!checked(hipEventElapsedTime(&ms, m.second.first, m.second.second), err, err_len, "AMR profile elapsed"))
        return 1;
      if (!totals.count(m.first)) order.push_back(m.first);
      totals[m.first].first += ms;
      totals[m.first].second += 1;
    }
    int32_t n = 0;
    for (const std::string& name : order) {
      if (n >= capacity) break;
      amr_kernel_time& t = out[n++];
      std::memset(t.name, 0, sizeof(t.name));
      std::strncpy(t.name, name.c_str(), sizeof(t.name) - 1);
      t.total_ms = totals[name].first;
      t.launches = totals[name].second;
    }
    *count = n;
    return 0;
  }
};

}  // namespace hipamr

#endif
