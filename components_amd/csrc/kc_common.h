// Shared definitions of libkacomp: error plumbing, device arena, kernel timing
// and wave-level helpers for gfx950 (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "kacomp.h"

namespace kc {

constexpr uint32_t kNone = 0xFFFFFFFFu;     // "no vertex" / empty hash slot
constexpr uint64_t kOneHop = 1ull << 32;    // +1 on the deviate half of a key
constexpr int kBlock = 256;                 // threads per workgroup (4 waves)
constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Errors: library code throws Failure after recording the message; the C ABI
// entry points catch it and return the code.
struct Failure {
  int code;
};
void set_error(const char *fmt, ...);
[[noreturn]] void fail(int code, const char *fmt, ...);

#define KC_HIP(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      ::kc::fail(KC_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                 __LINE__);                                                                 \
  } while (0)

#define KC_LAUNCH_CHECK() KC_HIP(hipGetLastError())

__host__ __device__ inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// Device arena: bump allocation out of chunks that are kept across calls, so
// that after the first call on a graph no device allocation happens.
class Arena {
 public:
  Arena() = default;
  Arena(const Arena &) = delete;
  Arena &operator=(const Arena &) = delete;
  ~Arena();
  void *alloc(size_t bytes);
  template <class T>
  T *alloc_n(size_t n) {
    return static_cast<T *>(alloc(n * sizeof(T) + 16));
  }
  void reset() {
    cur_ = 0;
    used_ = 0;
  }
  size_t reserved() const;

 private:
  std::vector<std::pair<char *, size_t>> chunks_;
  size_t cur_ = 0;   // chunk in use
  size_t used_ = 0;  // bytes used in that chunk
};

// ---------------------------------------------------------------------------
// Kernel timing with HIP events recorded on the launch stream.  Every launch
// group is one record (class, algorithmic bytes, optional event pair); a
// record can be re-tagged once its outcome is known (speculative sweeps that
// returned at once are moved to KC_K_OTHER with zero bytes).
class KernelTimer {
 public:
  ~KernelTimer();
  void begin_run(bool enabled, hipStream_t s);
  int start(int cls);
  void stop(int slot, uint64_t alg_bytes);
  void retag(int slot, int cls, uint64_t alg_bytes);
  // Synchronises (when timing) and fills the per-class fields of st.
  void finish(kc_stats *st);

 private:
  bool enabled_ = false;
  hipStream_t stream_ = nullptr;
  std::vector<hipEvent_t> pool_;
  struct Rec {
    int cls;
    int ev;
    uint64_t bytes;
  };
  std::vector<Rec> recs_;
  size_t next_ev_ = 0;
};

struct TimedScope {
  KernelTimer &t;
  int slot;
  uint64_t bytes;
  TimedScope(KernelTimer &tm, int cls, uint64_t b) : t(tm), slot(tm.start(cls)), bytes(b) {}
  ~TimedScope() { t.stop(slot, bytes); }
};

// ---------------------------------------------------------------------------
// Device helpers.
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int mask, int width) {
  return static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(v), mask, width));
}
__device__ __forceinline__ uint64_t min_u64(uint64_t a, uint64_t b) { return a < b ? a : b; }

// Min / sum over groups of `width` consecutive lanes (width a power of two <= 64).
__device__ __forceinline__ uint64_t group_min_u64(uint64_t v, int width) {
  for (int o = width >> 1; o > 0; o >>= 1) v = min_u64(v, shfl_xor_u64(v, o, width));
  return v;
}
__device__ __forceinline__ uint32_t group_sum_u32(uint32_t v, int width) {
  for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, width);
  return v;
}

}  // namespace kc
