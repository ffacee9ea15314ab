// Stable LSD radix sort of (u32 key, u32 value) pairs, 8-bit digits.
//
// Per pass: (1) every tile of 4096 items builds its digit histogram in LDS;
// (2) an exclusive scan over the digit-major histogram table gives each
// (digit, tile) its output offset; (3) every tile ranks its items in index
// order -- peers with the same digit inside a wave are found with 8 ballots,
// waves are ordered through LDS counts -- places them in LDS in digit order
// and writes each digit's items out as one contiguous run.  No global
// atomics; traffic per pass is 8 B read twice and 8 B written per item.
#include <hip/hip_runtime.h>

#include "kc_common.h"
#include "kernels.h"

namespace kc {

namespace {

constexpr int kRadixBits = 8;
constexpr int kDigits = 1 << kRadixBits;
constexpr int kItemsPerThread = 16;
constexpr int kTile = kBlock * kItemsPerThread;
constexpr int kWaves = kBlock / kWave;

__global__ __launch_bounds__(kBlock) void k_radix_hist(uint64_t n, const uint32_t *keys, int shift, uint32_t *hist,
                                                      uint32_t tiles) {
  __shared__ uint32_t cnt[kDigits];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kTile;
  for (int r = 0; r < kItemsPerThread; ++r) {
    const uint64_t i = base + static_cast<uint64_t>(r) * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (kDigits - 1)], 1u);
  }
  __syncthreads();
  hist[static_cast<uint64_t>(threadIdx.x) * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive prefix of one value per thread over the block (four waves).
__device__ __forceinline__ uint32_t block_exclusive_u32(uint32_t x, uint32_t *wave_tot) {
  const uint32_t lane = lane_id();
  uint32_t inc = x;
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(inc), o));
    if (lane >= static_cast<uint32_t>(o)) inc += y;
  }
  if (lane == kWave - 1) wave_tot[threadIdx.x / kWave] = inc;
  __syncthreads();
  uint32_t before = 0;
  for (int w = 0; w < static_cast<int>(threadIdx.x / kWave); ++w) before += wave_tot[w];
  __syncthreads();
  return before + inc - x;
}

// Scatter pass, staged in LDS: every wave owns a contiguous quarter of the
// tile and ranks its items on its own -- peers with the same digit inside a
// wave step found with 8 ballots, earlier steps through the wave's running
// per-digit counts in LDS -- so that one barrier orders the waves (per digit,
// the counts of the waves before) instead of barriers at every step.  Then
// the items are placed in LDS in digit order and consecutive threads write
// consecutive items, so each digit's items leave the tile as one contiguous
// run per array instead of one scattered 4-byte store each.
constexpr int kWaveItems = kWave * kItemsPerThread;  // a wave's share of the tile

__global__ __launch_bounds__(kBlock) void k_radix_scatter(uint64_t n, const uint32_t *keys, const uint32_t *vals,
                                                         int shift, const uint32_t *hist, const uint64_t *offsets,
                                                         uint32_t tiles, uint32_t *okeys, uint32_t *ovals) {
  __shared__ uint32_t skey[kTile];
  __shared__ uint32_t sval[kTile];
  __shared__ uint32_t local_start[kDigits];  // first tile position of each digit
  __shared__ uint32_t wave_cnt[kWaves][kDigits];  // per wave: running, then exclusive over waves
  __shared__ uint32_t wave_tot[kWaves];
  __shared__ uint64_t digit_base[kDigits];  // output offset of the tile's first item of each digit
  const int tid = threadIdx.x;
  const int w = tid / kWave;
  const uint32_t lane = lane_id();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kTile;
  const uint32_t count = static_cast<uint32_t>(n - base < static_cast<uint64_t>(kTile) ? n - base : kTile);
  // the wave's items come into registers first, all loads in flight at once:
  // step r, lane l holds tile item w * kWaveItems + r * kWave + l
  const uint32_t first = static_cast<uint32_t>(w) * kWaveItems + lane;
  uint32_t kr[kItemsPerThread], vr[kItemsPerThread];
#pragma unroll
  for (int r = 0; r < kItemsPerThread; ++r) {
    const uint32_t i = first + static_cast<uint32_t>(r) * kWave;
    kr[r] = i < count ? keys[base + i] : 0u;
    vr[r] = i < count ? vals[base + i] : 0u;
  }
  local_start[tid] = block_exclusive_u32(hist[static_cast<uint64_t>(tid) * tiles + blockIdx.x], wave_tot);
  digit_base[tid] = offsets[static_cast<uint64_t>(tid) * tiles + blockIdx.x];
  for (int d = lane; d < kDigits; d += kWave) wave_cnt[w][d] = 0;
  // (the wave reads only its own counts below, and a wave's LDS accesses
  // complete in issue order)
  __builtin_amdgcn_wave_barrier();
  uint32_t rank[kItemsPerThread];
#pragma unroll
  for (int r = 0; r < kItemsPerThread; ++r) {
    const bool valid = first + static_cast<uint32_t>(r) * kWave < count;
    const uint32_t d = (kr[r] >> shift) & (kDigits - 1);
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < kRadixBits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t below = static_cast<uint32_t>(__popcll(peers & lanes_below()));
    const uint32_t before = valid ? wave_cnt[w][d] : 0u;
    __builtin_amdgcn_wave_barrier();
    rank[r] = before + below;
    if (valid && below == 0) wave_cnt[w][d] = before + static_cast<uint32_t>(__popcll(peers));
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  {  // per digit, the counts of the waves before
    uint32_t acc = 0;
    for (int w2 = 0; w2 < kWaves; ++w2) {
      const uint32_t c = wave_cnt[w2][tid];
      wave_cnt[w2][tid] = acc;
      acc += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kItemsPerThread; ++r) {
    if (first + static_cast<uint32_t>(r) * kWave < count) {
      const uint32_t d = (kr[r] >> shift) & (kDigits - 1);
      const uint32_t pos = local_start[d] + wave_cnt[w][d] + rank[r];
      skey[pos] = kr[r];
      sval[pos] = vr[r];
    }
  }
  __syncthreads();
  for (uint32_t j = tid; j < count; j += kBlock) {
    const uint32_t k = skey[j];
    const uint32_t d = (k >> shift) & (kDigits - 1);
    const uint64_t g = digit_base[d] + (j - local_start[d]);
    okeys[g] = k;
    ovals[g] = sval[j];
  }
}

template <class T>
struct ArrLoad {
  const T *p;
  __device__ uint64_t operator()(uint64_t i) const { return p[i]; }
};
struct ArrStore64 {
  uint64_t *p;
  __device__ void operator()(uint64_t i, uint64_t prefix, uint64_t) const { p[i] = prefix; }
};

}  // namespace

uint64_t radix_sort_scratch_words(uint64_t n) {
  const uint64_t tiles = ceil_div(n, kTile) + 1;
  return tiles * kDigits * 2 + scan_scratch_words(tiles * kDigits) + 64;
}

// Sorts pairs by the low `bits` bits of the key.  Returns the buffer pair
// (0: keys/vals, 1: tkeys/tvals) that holds the result.
int radix_sort_pairs(uint64_t n, int bits, uint32_t *keys, uint32_t *vals, uint32_t *tkeys, uint32_t *tvals,
                     uint64_t *scratch, hipStream_t s) {
  if (n <= 1 || bits <= 0) return 0;
  const uint32_t tiles = static_cast<uint32_t>(ceil_div(n, kTile));
  const uint64_t cells = static_cast<uint64_t>(tiles) * kDigits;
  uint32_t *hist = reinterpret_cast<uint32_t *>(scratch);
  uint64_t *offs = scratch + ceil_div(cells, 2) + 8;
  uint64_t *total = offs + cells + 8;
  uint64_t *rest = total + 8;
  int cur = 0;
  for (int shift = 0; shift < bits; shift += kRadixBits) {
    uint32_t *ik = cur ? tkeys : keys, *iv = cur ? tvals : vals;
    uint32_t *ok = cur ? keys : tkeys, *ov = cur ? vals : tvals;
    hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kBlock), 0, s, n, ik, shift, hist, tiles);
    KC_LAUNCH_CHECK();
    scan_exclusive_u32_to_u64(cells, hist, offs, total, rest, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kBlock), 0, s, n, ik, iv, shift, hist, offs, tiles, ok,
                       ov);
    KC_LAUNCH_CHECK();
    cur ^= 1;
  }
  return cur;
}

}  // namespace kc
