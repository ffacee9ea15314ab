// kc_find_components_noalloc: the propagation variants of kc_find_components
// (KC_VARIANT_LABELS, min-label propagation, Propagation::FindComponents,
// lib/components/propagation.h:48-53; KC_VARIANT_SHORTCUT, propagation with
// pointer jumping, ShortcutPropagation::FindComponents,
// lib/components/shortcut_propagation.h:89-94) on a CSR the caller already
// holds in device memory, without any device or pinned host allocation.
//
// Every work buffer of those variants (keys, degree bins, hub chunks, sweep
// flags, counters, partial sums, the binning scan's scratch) is carved out of
// one file-scope __device__ array, `scratch`, sized for graphs of up to
// KC_NOALLOC_MAX_VERTICES vertices and KC_NOALLOC_MAX_EDGE_SLOTS edge slots.
// Each GPU holds its own copy of the array for the life of the process; its
// device address is looked up once per device (hipGetSymbolAddress) and only
// that address is handed to kernels.  The kernels are the ones
// kc_find_components launches for these variants, in the same schedule, so the
// labels are the same (the number of sweeps can differ between any two runs:
// in-place sweeps read keys that other threads are lowering).
//
// Every call on one GPU uses that GPU's array: calls that overlap on one GPU
// (several host threads, or several streams in flight) race and must be
// serialised by the caller.  Whatever must start cleared (the sweep flags) is
// cleared by a fill kernel on every call.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <mutex>

#include "kacomp.h"
#include "kc_common.h"
#include "kernels.h"

namespace kc {
namespace {

constexpr uint64_t kMaxVertices = KC_NOALLOC_MAX_VERTICES;
constexpr uint64_t kMaxSlots = KC_NOALLOC_MAX_EDGE_SLOTS;
constexpr uint32_t kFlagSlots = 1024;  // sweep flags, reused round-robin
constexpr uint64_t kBinScratchBytes = 1ull << 20;

constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }
constexpr uint64_t kHubCap = kMaxSlots / (kMinWaveBin / 2) + 2;  // hub_chunk_cap(kMaxSlots)

// Byte offsets of the regions inside `scratch`.
constexpr uint64_t kKeyAt = 0;                                                  // u64[n]
constexpr uint64_t kListsAt = align256(kKeyAt + 8 * kMaxVertices);              // u32[3n]
constexpr uint64_t kDegAt = align256(kListsAt + 12 * kMaxVertices);             // u32[n]
constexpr uint64_t kHubsAt = align256(kDegAt + 4 * kMaxVertices);               // HubChunk[kHubCap]
constexpr uint64_t kHubInterAt = align256(kHubsAt + sizeof(HubChunk) * kHubCap);  // u32[kHubCap]
constexpr uint64_t kFlagsAt = align256(kHubInterAt + 4 * kHubCap);              // int[kFlagSlots]
constexpr uint64_t kSmallAt = align256(kFlagsAt + 4 * kFlagSlots);              // u32[16]: bin counts, roots
constexpr uint64_t kPartialAt = align256(kSmallAt + 64);                        // u32[n/256 + 64]
constexpr uint64_t kBinScratchAt = align256(kPartialAt + 4 * (kMaxVertices / kBlock + 64));
constexpr uint64_t kScratchBytes = align256(kBinScratchAt + kBinScratchBytes);

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint8_t *g_scratch_at[kMaxDevices] = {};

// The device address of this device's `scratch` (looked up on the first call
// on the device, with the device current).
uint8_t *device_scratch() {
  int dev = 0;
  KC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(KC_ERR_UNSUPPORTED, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_at[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      fail(KC_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_at[dev] = static_cast<uint8_t *>(p);
  }
  return g_scratch_at[dev];
}

template <class T>
T *region(uint8_t *base, uint64_t at) {
  return reinterpret_cast<T *>(base + at);
}

}  // namespace
}  // namespace kc

using namespace kc;

extern "C" int kc_find_components_noalloc(uint64_t n, const uint64_t *offsets, const uint32_t *targets,
                                          const kc_options *opt, uint32_t *labels, void *stream, kc_stats *stats) {
  return guarded([&] {
    kc_options o;
    if (opt)
      o = *opt;
    else
      kc_options_init(&o);
    if (o.variant != KC_VARIANT_LABELS && o.variant != KC_VARIANT_SHORTCUT)
      fail(KC_ERR_UNSUPPORTED, "kc_find_components_noalloc runs the labels and shortcuts variants only (got %d)",
           o.variant);
    if (n > kMaxVertices)
      fail(KC_ERR_INVALID, "kc_find_components_noalloc: n=%llu exceeds the scratch limit of %llu vertices",
           static_cast<unsigned long long>(n), static_cast<unsigned long long>(kMaxVertices));
    if (n && (!offsets || !labels)) fail(KC_ERR_INVALID, "null device pointer");
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint64_t ends[2] = {0, 0};  // offsets[0], offsets[n]
    if (n) {
      KC_HIP(hipMemcpyAsync(&ends[0], offsets, 8, hipMemcpyDeviceToHost, s));
      KC_HIP(hipMemcpyAsync(&ends[1], offsets + n, 8, hipMemcpyDeviceToHost, s));
      KC_HIP(hipStreamSynchronize(s));
    }
    const uint64_t m2 = ends[1];
    if (ends[0] != 0) fail(KC_ERR_INVALID, "offsets[0] must be 0");
    if (m2 > kMaxSlots)
      fail(KC_ERR_INVALID, "kc_find_components_noalloc: 2m=%llu exceeds the scratch limit of %llu edge slots",
           static_cast<unsigned long long>(m2), static_cast<unsigned long long>(kMaxSlots));
    if (m2 && !targets) fail(KC_ERR_INVALID, "null device pointer");
    if (bin_scratch_words(n) * 8 > kBinScratchBytes) fail(KC_ERR_INVALID, "internal: binning scratch too small");
    uint64_t sweeps = 0;
    uint32_t components = 0;
    if (n) {
      uint8_t *base = device_scratch();
      uint64_t *key = region<uint64_t>(base, kKeyAt);
      uint32_t *lists = region<uint32_t>(base, kListsAt);
      uint32_t *deg_inter = region<uint32_t>(base, kDegAt);
      HubChunk *hubs = region<HubChunk>(base, kHubsAt);
      uint32_t *hub_inter = region<uint32_t>(base, kHubInterAt);
      int *flags = region<int>(base, kFlagsAt);
      uint32_t *small = region<uint32_t>(base, kSmallAt);
      uint32_t *partial = region<uint32_t>(base, kPartialAt);
      uint64_t *bin_scratch = region<uint64_t>(base, kBinScratchAt);
      // degree bins of the graph (as make_bins does for these variants)
      const BinCfg cfg = bin_config(o.sweep_group_small, o.sweep_group_mid, o.sweep_group_large);
      Bins b;
      b.listT = lists;
      b.listS = lists + n;
      b.listW = lists + 2 * n;
      b.hubs = hubs;
      bin_vertices(n, offsets, cfg, b.listT, b.listS, b.listW, b.hubs, small, deg_inter, bin_scratch, s);
      uint32_t counts[4];
      KC_HIP(hipMemcpyAsync(counts, small, sizeof(counts), hipMemcpyDeviceToHost, s));
      KC_HIP(hipStreamSynchronize(s));
      b.nT = counts[0];
      b.nS = counts[1];
      b.nW = counts[2];
      b.nH = counts[3];
      if (b.nH > kHubCap) fail(KC_ERR_INVALID, "internal: hub chunk overflow (%u)", b.nH);
      finish_bins(b);
      init_keys_index(n, key, s);
      // sweeps to the fixed point in speculative batches, as kc_find_components
      // runs them: a sweep whose predecessor changed nothing returns at once
      const bool with_jump = o.variant == KC_VARIANT_SHORTCUT;
      const uint32_t batch = n < (1u << 16) ? 8u : 3u;
      uint32_t flag_base = 0;
      fill_u32(reinterpret_cast<uint32_t *>(flags), kFlagSlots, 0, s);
      while (true) {
        if (flag_base + batch + 2 >= kFlagSlots) {
          fill_u32(reinterpret_cast<uint32_t *>(flags), kFlagSlots, 0, s);
          flag_base = 0;
        }
        for (uint32_t i = 0; i < batch; ++i) {
          const uint32_t f = flag_base + i;
          const int *prev = i == 0 ? nullptr : &flags[f - 1];
          sweep(kSweepInPlace, false, offsets, targets, key, key, 0, b, deg_inter, hub_inter, &flags[f], prev, s);
          if (with_jump) jump(n, key, &flags[f], s);
        }
        int h_flags[8];
        KC_HIP(hipMemcpyAsync(h_flags, &flags[flag_base], batch * sizeof(int), hipMemcpyDeviceToHost, s));
        KC_HIP(hipStreamSynchronize(s));
        uint32_t i = 0;
        while (i < batch && h_flags[i] != 0) ++i;
        sweeps += i < batch ? i + 1 : batch;
        flag_base += batch;
        if (i < batch) break;
      }
      labels_from_keys(n, key, labels, small + 4, partial, s);
      KC_HIP(hipMemcpyAsync(&components, small + 4, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      KC_HIP(hipStreamSynchronize(s));
    }
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      stats->n = n;
      stats->m = m2 / 2;
      stats->components = components;
      stats->sweeps = sweeps;
    }
  });
}
