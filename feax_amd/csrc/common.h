// Shared helpers of libfeaxhip: error reporting and launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "feaxhip.h"

namespace fx {

// Per-thread message of the last failure (fx_last_error).
inline std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}

inline int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  last_error() = buf;
  return code;
}

inline int hip_fail(hipError_t e, const char* where) {
  return fail(static_cast<int>(e), "%s: %s", where, hipGetErrorString(e));
}

#define FX_HIP(call)                                  \
  do {                                                \
    hipError_t fx_e_ = (call);                        \
    if (fx_e_ != hipSuccess) return fx::hip_fail(fx_e_, #call); \
  } while (0)

#define FX_LAUNCHED(name)                             \
  do {                                                \
    hipError_t fx_e_ = hipGetLastError();             \
    if (fx_e_ != hipSuccess) return fx::hip_fail(fx_e_, name); \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Owns a captured HIP graph, its executable and the private capture stream.
struct GraphGuard {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t stream = nullptr;
  GraphGuard() = default;
  GraphGuard(const GraphGuard&) = delete;
  GraphGuard& operator=(const GraphGuard&) = delete;
  ~GraphGuard() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Owns HIP events.
struct EventGuard {
  std::vector<hipEvent_t> ev;
  EventGuard() = default;
  EventGuard(const EventGuard&) = delete;
  EventGuard& operator=(const EventGuard&) = delete;
  ~EventGuard() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

// MI355X: 256 CUs in 8 XCDs; workgroups are dealt to XCDs round-robin.
constexpr int kNumXcd = 8;

// Map the hardware block id so that each XCD works on one contiguous range
// of logical blocks (keeps gathered vector entries in that XCD's L2).
__host__ __device__ __forceinline__ int xcd_block(int bid, int nblocks) {
  int per = (nblocks + kNumXcd - 1) / kNumXcd;
  int xcd = bid % kNumXcd;
  int idx = bid / kNumXcd;
  int full = nblocks - (per - 1) * kNumXcd;  // XCDs that receive `per` blocks
  // XCD x receives blocks x, x+8, ...: `per` of them if x < full, else per-1.
  int base = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
  return base + idx;
}

}  // namespace fx
