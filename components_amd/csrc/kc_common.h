// Shared definitions of libkacomp: error plumbing, device arena, kernel timing
// and wave-level helpers for gfx950 (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <new>
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

// Runs f and turns a Failure (or any exception) into the C ABI's return code.
template <class F>
int guarded(F &&f) {
  try {
    f();
    return KC_OK;
  } catch (const Failure &e) {
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return KC_ERR_OOM;
  } catch (...) {
    set_error("unexpected exception");
    return KC_ERR_INVALID;
  }
}

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
// returned at once keep their class with zero bytes).
class KernelTimer {
 public:
  ~KernelTimer();
  // only_cls >= 0: events for the launches of that class alone.
  void begin_run(bool enabled, hipStream_t s, int only_cls = -1);
  int start(int cls);
  void stop(int slot, uint64_t alg_bytes);
  void retag(int slot, int cls, uint64_t alg_bytes);
  // Synchronises (when timing) and fills the per-class fields of st.
  void finish(kc_stats *st);

 private:
  bool enabled_ = false;
  int only_ = -1;
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

// Atomic fetch-add on arr[key] for every lane with part set.  The lanes that
// share the key of the first participating lane are merged into one atomic
// (hub clusters and giant components make many lanes of a wave hit the same
// counter); the remaining lanes issue their own.  Returns the old value each
// lane would have seen with one atomic per lane in lane order.
__device__ __forceinline__ uint32_t peeled_fetch_add(uint32_t *arr, uint32_t key, uint32_t val, bool part) {
  const uint64_t pm = __ballot(part);
  if (pm == 0) return 0;
  const int leader = __ffsll(static_cast<unsigned long long>(pm)) - 1;
  const uint32_t lk = __shfl(key, leader);
  const bool same = part && key == lk;
  const uint32_t x = same ? val : 0u;
  uint32_t incl = x;
  const int ln = static_cast<int>(lane_id());
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o);
    if (ln >= o) incl += t;
  }
  const uint32_t total = __shfl(incl, 63);
  uint32_t base = 0;
  if (ln == leader) base = atomicAdd(&arr[lk], total);
  base = __shfl(base, leader);
  if (same) return base + incl - x;
  if (part) return atomicAdd(&arr[key], val);
  return 0;
}

// atomicMin(&arr[key], val) for every lane with part set, merged for the
// lanes sharing the first lane's key and skipped when arr[key] is already
// smaller (values only decrease, so a stale read only costs an atomic).
__device__ __forceinline__ void peeled_min(uint32_t *arr, uint32_t key, uint32_t val, bool part) {
  const uint64_t pm = __ballot(part);
  if (pm == 0) return;
  const int leader = __ffsll(static_cast<unsigned long long>(pm)) - 1;
  const uint32_t lk = __shfl(key, leader);
  const bool same = part && key == lk;
  uint32_t mn = same ? val : 0xFFFFFFFFu;
  for (int o = 32; o > 0; o >>= 1) mn = min(mn, static_cast<uint32_t>(__shfl_xor(mn, o)));
  if (static_cast<int>(lane_id()) == leader && mn < arr[lk]) atomicMin(&arr[lk], mn);
  if (part && !same && val < arr[key]) atomicMin(&arr[key], val);
}

// Plain-path handles built from a CSR already in device memory (used by the
// batched form for the union of its large graphs).  graph_device_begin
// allocates the handle and its level-0 arrays (n + 1 offsets, m2 columns),
// which the caller fills on the returned stream; graph_device_finish then
// counts the non-isolated vertices and allocates the work buffers.  The
// handle is released with kc_graph_destroy.
kc_graph *graph_device_begin(uint64_t n, uint64_t m2, uint64_t **off, uint32_t **col, hipStream_t *stream);
void graph_device_finish(kc_graph *g);
// Device labels of the handle's last kc_find_components call.
const uint32_t *graph_device_labels(const kc_graph *g);

}  // namespace kc
