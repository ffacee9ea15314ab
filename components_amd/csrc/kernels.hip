// Device kernels of the contraction path for gfx950 (wave64, 160 KB LDS/CU).
//
// Data layout per level l (all in HBM):
//   off   u64[n_l+1]  CSR row offsets        col  u32[2 m_l]  neighbour indices
//   key   u64[n_l]    (deviate + hops) << 32 | centre index  (lexicographic key
//                     of lib/components/exponential_contraction.h:531-537 as one integer)
//   lab   u32[n_l]    centre index after the decomposition
//   cnt   u32[n_l]    cluster-edge count of each centre's bucket
//   bucket u32[P]     next-level neighbour indices grouped by cluster (P = sum cnt)
// Integer work only; no MFMA.  Every kernel is bandwidth-bound.
#include <hip/hip_runtime.h>

#include "deviates.h"
#include "kc_common.h"
#include "kernels.h"

namespace kc {

namespace {

inline unsigned grid_for(uint64_t n, unsigned per_block = kBlock) {
  uint64_t g = ceil_div(n, per_block);
  if (g == 0) g = 1;
  if (g > 0x7fffffffull) g = 0x7fffffffull;
  return static_cast<unsigned>(g);
}

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Wave-aggregated append: every lane with pred gets a distinct slot of *counter.
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t *counter) {
  uint64_t m = __ballot(pred);
  if (m == 0) return 0;
  int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
  uint32_t base = 0;
  if (static_cast<int>(lane_id()) == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(m)));
  base = __shfl(base, leader);
  return base + static_cast<uint32_t>(__popcll(m & lanes_below()));
}

// ---------------------------------------------------------------------------
// Fills.
template <class T>
__global__ void k_fill(T *p, uint64_t n, T v) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// Keys.
// Key of the vertex stored at v: (deviate << 32) | level index, where the
// level index is idx[v] when the level is stored in another order (null: v).
__global__ void k_init_keys_exp(uint64_t n, const uint32_t *gid, const uint32_t *idx, DeviateArgs d, LearnedArgs la,
                                uint64_t *key) {
  uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (v >= n) return;
  uint32_t g = gid ? gid[v] : static_cast<uint32_t>(v);
  uint32_t dev = exp_deviate(g, d.seed, d.iteration, d.rng_offset, d.fused != 0, d.stream, d.table, d.k_max);
  if (la.delay) dev += la.delay[v];
  key[v] = (static_cast<uint64_t>(dev) << 32) | (idx ? idx[v] : static_cast<uint32_t>(v));
}

__global__ void k_init_keys_index(uint64_t n, uint64_t *key) {
  uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (v < n) key[v] = static_cast<uint32_t>(v);
}

__global__ void k_init_keys_uniform(uint64_t n, const uint32_t *gid, const uint32_t *idx, int32_t seed, uint32_t it,
                                    uint64_t *key) {
  uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (v >= n) return;
  uint32_t g = gid ? gid[v] : static_cast<uint32_t>(v);
  key[v] = (static_cast<uint64_t>(uniform_deviate_100(g, seed, it)) << 32) | (idx ? idx[v] : static_cast<uint32_t>(v));
}

__global__ void k_deviates_only(uint32_t gid0, uint32_t count, DeviateArgs d, uint32_t *out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  out[i] = exp_deviate(gid0 + i, d.seed, d.iteration, d.rng_offset, d.fused != 0, d.stream, d.table, d.k_max);
}

// ---------------------------------------------------------------------------
// Degree binning, two passes without global atomics: per-block bin counts,
// one exclusive scan over the bin-major count table, then a stable write.
__device__ __forceinline__ int bin_of(uint64_t deg, const BinCfg &cfg) {
  return deg == 0 ? -1 : deg <= cfg.tmax ? 0 : deg <= cfg.smax ? 1 : deg <= cfg.wmax ? 2 : 3;
}

// Per wave: rank of each lane among the lanes of its bin, and the wave's
// per-bin totals in LDS (bins 0..2 count vertices, bin 3 counts hub chunks).
__device__ __forceinline__ void bin_ranks(int bin, uint32_t chunks, uint32_t *rank, uint32_t wave_tot[4]) {
  for (int r = 0; r < 3; ++r) {
    const uint64_t m = __ballot(bin == r);
    if (bin == r) *rank = static_cast<uint32_t>(__popcll(m & lanes_below()));
    wave_tot[r] = static_cast<uint32_t>(__popcll(m));
  }
  uint32_t x = bin == 3 ? chunks : 0u, incl = x;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o);
    if (static_cast<int>(lane_id()) >= o) incl += t;
  }
  if (bin == 3) *rank = incl - x;
  wave_tot[3] = __shfl(incl, 63);
}

// With key != null only vertices at key distance >= tmin are binned (the
// unsettled ones of the time-stepped schedule).
__device__ __forceinline__ uint64_t binned_degree(uint64_t v, uint64_t n, const uint64_t *off, const uint64_t *key,
                                                  uint32_t tmin) {
  if (v >= n) return 0;
  if (key && static_cast<uint32_t>(key[v] >> 32) < tmin) return 0;
  return off[v + 1] - off[v];
}

__global__ __launch_bounds__(kBlock) void k_bin_count(uint64_t n, const uint64_t *off, BinCfg cfg, uint32_t *counts,
                                                     uint32_t nblocks, const uint64_t *key, uint32_t tmin) {
  __shared__ uint32_t tot[kBlock / kWave][4];
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
  const uint64_t deg = binned_degree(v, n, off, key, tmin);
  const int bin = bin_of(deg, cfg);
  const uint32_t chunks = bin == 3 ? static_cast<uint32_t>(ceil_div(deg, cfg.wmax)) : 0u;
  uint32_t rank = 0, wt[4];
  bin_ranks(bin, chunks, &rank, wt);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0)
    for (int r = 0; r < 4; ++r) tot[w][r] = wt[r];
  __syncthreads();
  if (threadIdx.x < 4)
    counts[static_cast<uint64_t>(threadIdx.x) * nblocks + blockIdx.x] =
        tot[0][threadIdx.x] + tot[1][threadIdx.x] + tot[2][threadIdx.x] + tot[3][threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_bin_write(uint64_t n, const uint64_t *off, BinCfg cfg,
                                                     const uint64_t *prefix, uint32_t nblocks, uint32_t *listT,
                                                     uint32_t *listS, uint32_t *listW, HubChunk *hubs,
                                                     uint32_t *deg_inter, const uint64_t *key, uint32_t tmin) {
  __shared__ uint32_t tot[kBlock / kWave][4];
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
  const uint64_t deg = binned_degree(v, n, off, key, tmin);
  const uint64_t e0 = deg ? off[v] : 0;
  if (v < n && deg == 0 && deg_inter) deg_inter[v] = 0;
  const int bin = bin_of(deg, cfg);
  const uint32_t chunks = bin == 3 ? static_cast<uint32_t>(ceil_div(deg, cfg.wmax)) : 0u;
  uint32_t rank = 0, wt[4];
  bin_ranks(bin, chunks, &rank, wt);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0)
    for (int r = 0; r < 4; ++r) tot[w][r] = wt[r];
  __syncthreads();
  if (bin < 0) return;
  uint64_t pos = prefix[static_cast<uint64_t>(bin) * nblocks + blockIdx.x] - prefix[static_cast<uint64_t>(bin) * nblocks];
  for (int w2 = 0; w2 < w; ++w2) pos += tot[w2][bin];
  pos += rank;
  if (bin == 0) {
    listT[pos] = static_cast<uint32_t>(v);
  } else if (bin == 1) {
    listS[pos] = static_cast<uint32_t>(v);
  } else if (bin == 2) {
    listW[pos] = static_cast<uint32_t>(v);
  } else {
    for (uint32_t c = 0; c < chunks; ++c) {
      HubChunk h;
      h.v = static_cast<uint32_t>(v);
      h.k = c;
      h.e0 = e0 + static_cast<uint64_t>(c) * cfg.wmax;
      h.e1 = min(e0 + deg, h.e0 + cfg.wmax);
      hubs[pos + c] = h;
    }
  }
}

__global__ void k_bin_totals(const uint64_t *prefix, const uint64_t *total, uint32_t nblocks, uint32_t *counters) {
  const int r = threadIdx.x;
  if (r >= 4) return;
  const uint64_t hi = r < 3 ? prefix[static_cast<uint64_t>(r + 1) * nblocks] : *total;
  counters[r] = static_cast<uint32_t>(hi - prefix[static_cast<uint64_t>(r) * nblocks]);
}

// ---------------------------------------------------------------------------
// Binned sweep.  The block index selects the bin section; inside a section a
// group of W lanes owns one vertex (or one hub chunk) and strides over its
// edges, then reduces with cross-lane shuffles.
struct SweepArgs {
  const uint64_t *off;
  const uint32_t *col;
  const uint64_t *kin;
  uint64_t *kout;
  uint64_t hop;
  Bins b;
  uint32_t *deg_inter;
  uint32_t *hub_inter;
  int *flag;
  const int *prev_flag;
  uint64_t block = 0;  // kSweepBlock: indices per block
  int hubs_first = 0;  // k_sweep: sections dispatched from the hub chunks down (as k_round's)
};

// Edges whose gathers one lane keeps in flight together in the row loops.
constexpr int kGather = 4;
constexpr int kRelabelGather = 16;

template <int MODE, bool COUNT, int W>
__device__ __forceinline__ bool sweep_group(const SweepArgs &a, bool active, uint32_t v, uint64_t e0,
                                            uint64_t e1, int gl, bool hub, uint32_t chunk) {
  uint64_t kv = active ? a.kin[v] : ~0ull;
  uint64_t best = kv;
  uint32_t cnt = 0;
  if (active) {
    const uint32_t lv = static_cast<uint32_t>(kv);
    // kSweepBlock reads the keys of the neighbours in v's block only
    const uint64_t lo = MODE == kSweepBlock ? v / a.block * a.block : 0;
    const uint64_t hi = MODE == kSweepBlock ? lo + a.block : ~0ull;
    uint64_t e = e0 + gl;
    // kGather edges per lane per step: their neighbour ids, then their keys,
    // are all in flight together
    for (; e + (kGather - 1) * W < e1; e += kGather * W) {
      uint32_t c[kGather];
      uint64_t k[kGather];
#pragma unroll
      for (int j = 0; j < kGather; ++j) c[j] = a.col[e + j * W];
#pragma unroll
      for (int j = 0; j < kGather; ++j) k[j] = c[j] >= lo && c[j] < hi ? a.kin[c[j]] : ~0ull;
#pragma unroll
      for (int j = 0; j < kGather; ++j) {
        best = min_u64(best, k[j] + a.hop);
        if (COUNT) cnt += static_cast<uint32_t>(k[j]) != lv;
      }
    }
    for (; e < e1; e += W) {
      const uint32_t c = a.col[e];
      const uint64_t kw = c >= lo && c < hi ? a.kin[c] : ~0ull;
      best = min_u64(best, kw + a.hop);
      if (COUNT) cnt += static_cast<uint32_t>(kw) != lv;
    }
  }
  if (W > 1) {
    best = group_min_u64(best, W);
    if (COUNT) cnt = group_sum_u32(cnt, W);
  }
  bool changed = false;
  if (active && gl == 0) {
    if (MODE != kSweepCountOnly && best < kv) {
      if (hub) {
        uint64_t old = atomicMin(reinterpret_cast<unsigned long long *>(&a.kout[v]),
                                 static_cast<unsigned long long>(best));
        changed = old > best;
      } else {
        a.kout[v] = best;
        changed = true;
      }
    }
    if (COUNT) {
      if (hub)
        a.hub_inter[chunk] = cnt;
      else
        a.deg_inter[v] = cnt;
    }
  }
  return changed;
}

template <int MODE, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_sweep(SweepArgs a) {
  if (a.prev_flag && *a.prev_flag == 0) return;
  uint32_t blk = blockIdx.x;
  if (a.hubs_first) {
    const uint32_t nh = a.b.blocks - a.b.bH, nw = a.b.bH - a.b.bW, ns = a.b.bW - a.b.bS;
    blk = blk < nh ? a.b.bH + blk
                   : blk < nh + nw ? a.b.bW + (blk - nh) : blk < nh + nw + ns ? a.b.bS + (blk - nh - nw) : blk - nh - nw - ns;
  }
  const int tid = threadIdx.x;
  bool changed = false;
  if (blk < a.b.bS) {
    const uint32_t i = blk * kBlock + tid;
    const bool act = i < a.b.nT;
    uint32_t v = act ? a.b.listT[i] : 0;
    uint64_t e0 = act ? a.off[v] : 0, e1 = act ? a.off[v + 1] : 0;
    changed = sweep_group<MODE, COUNT, 1>(a, act, v, e0, e1, 0, false, 0);
  } else if (blk < a.b.bW) {
    const uint32_t i = (blk - a.b.bS) * (kBlock / 16) + tid / 16;
    const bool act = i < a.b.nS;
    uint32_t v = act ? a.b.listS[i] : 0;
    uint64_t e0 = act ? a.off[v] : 0, e1 = act ? a.off[v + 1] : 0;
    changed = sweep_group<MODE, COUNT, 16>(a, act, v, e0, e1, tid & 15, false, 0);
  } else if (blk < a.b.bH) {
    const uint32_t i = (blk - a.b.bW) * (kBlock / kWave) + tid / kWave;
    const bool act = i < a.b.nW;
    uint32_t v = act ? a.b.listW[i] : 0;
    uint64_t e0 = act ? a.off[v] : 0, e1 = act ? a.off[v + 1] : 0;
    changed = sweep_group<MODE, COUNT, kWave>(a, act, v, e0, e1, tid & 63, false, 0);
  } else {
    const uint32_t i = (blk - a.b.bH) * (kBlock / kWave) + tid / kWave;
    const bool act = i < a.b.nH;
    HubChunk h = act ? a.b.hubs[i] : HubChunk{0, 0, 0, 0};
    changed = sweep_group<MODE, COUNT, kWave>(a, act, h.v, h.e0, h.e1, tid & 63, true, i);
  }
  if (MODE != kSweepCountOnly) {
    if (__ballot(changed) != 0 && lane_id() == 0) *a.flag = 1;
  }
}

// Packed keys for the pull sweeps of levels below 2^24 vertices whose key
// distances stay below 255: key (distance, label) as distance << 24 | label,
// the same order in half the bytes, so that a 128-byte line brings 32 keys to
// the gathers instead of 16.  *bad receives a nonzero value when a key does
// not fit (the caller then sweeps the 64-bit keys).
__global__ __launch_bounds__(kBlock) void k_pack_keys(uint64_t n, const uint64_t *key, uint32_t *pk, uint32_t *bad) {
  bool over = false;
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const uint64_t k = key[v];
    const uint32_t d = static_cast<uint32_t>(k >> 32), l = static_cast<uint32_t>(k);
    // a neighbour's key plus one hop must fit as well: distances up to 254
    const bool fits = d < kPackedMaxDist && l < (1u << kPackedLabelBits);
    over |= !fits;
    pk[v] = fits ? (d << kPackedLabelBits) | l : ~0u;
  }
  if (__ballot(over) != 0 && lane_id() == 0) *bad = 1;
}

__global__ void k_unpack_keys(uint64_t n, const uint32_t *pk, uint64_t *key) {
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (v >= n) return;
  const uint32_t k = pk[v];
  key[v] = (static_cast<uint64_t>(k >> kPackedLabelBits) << 32) | (k & ((1u << kPackedLabelBits) - 1u));
}

// The in-place pull sweep over packed keys (k_sweep's kSweepInPlace without
// counts): the same sections, section order and kGather loads in flight.
struct PackedSweepArgs {
  const uint64_t *off;
  const uint32_t *col;
  uint32_t *key;
  Bins b;
  int *flag;
  const int *prev_flag;
  int hubs_first;
};

template <int W>
__device__ __forceinline__ bool packed_group(const PackedSweepArgs &a, bool active, uint32_t v, uint64_t e0,
                                             uint64_t e1, int gl, bool hub) {
  const uint32_t kv = active ? a.key[v] : ~0u;
  uint32_t best = kv;
  if (active) {
    uint64_t e = e0 + gl;
    for (; e + (kGather - 1) * W < e1; e += kGather * W) {
      uint32_t c[kGather], k[kGather];
#pragma unroll
      for (int j = 0; j < kGather; ++j) c[j] = a.col[e + j * W];
#pragma unroll
      for (int j = 0; j < kGather; ++j) k[j] = a.key[c[j]];
#pragma unroll
      for (int j = 0; j < kGather; ++j) best = min(best, k[j] + kPackedHop);
    }
    for (; e < e1; e += W) best = min(best, a.key[a.col[e]] + kPackedHop);
  }
  if (W > 1)
    for (int o = W >> 1; o > 0; o >>= 1) best = min(best, static_cast<uint32_t>(__shfl_xor(best, o, W)));
  bool changed = false;
  if (active && gl == 0 && best < kv) {
    if (hub) {
      changed = atomicMin(&a.key[v], best) > best;
    } else {
      a.key[v] = best;
      changed = true;
    }
  }
  return changed;
}

__global__ __launch_bounds__(kBlock) void k_sweep_packed(PackedSweepArgs a) {
  if (a.prev_flag && *a.prev_flag == 0) return;
  uint32_t blk = blockIdx.x;
  if (a.hubs_first) {
    const uint32_t nh = a.b.blocks - a.b.bH, nw = a.b.bH - a.b.bW, ns = a.b.bW - a.b.bS;
    blk = blk < nh ? a.b.bH + blk
                   : blk < nh + nw ? a.b.bW + (blk - nh) : blk < nh + nw + ns ? a.b.bS + (blk - nh - nw) : blk - nh - nw - ns;
  }
  const int tid = threadIdx.x;
  bool changed = false;
  if (blk < a.b.bS) {
    const uint32_t i = blk * kBlock + tid;
    const bool act = i < a.b.nT;
    const uint32_t v = act ? a.b.listT[i] : 0;
    changed = packed_group<1>(a, act, v, act ? a.off[v] : 0, act ? a.off[v + 1] : 0, 0, false);
  } else if (blk < a.b.bW) {
    const uint32_t i = (blk - a.b.bS) * (kBlock / 16) + tid / 16;
    const bool act = i < a.b.nS;
    const uint32_t v = act ? a.b.listS[i] : 0;
    changed = packed_group<16>(a, act, v, act ? a.off[v] : 0, act ? a.off[v + 1] : 0, tid & 15, false);
  } else if (blk < a.b.bH) {
    const uint32_t i = (blk - a.b.bW) * (kBlock / kWave) + tid / kWave;
    const bool act = i < a.b.nW;
    const uint32_t v = act ? a.b.listW[i] : 0;
    changed = packed_group<kWave>(a, act, v, act ? a.off[v] : 0, act ? a.off[v + 1] : 0, tid & 63, false);
  } else {
    const uint32_t i = (blk - a.b.bH) * (kBlock / kWave) + tid / kWave;
    const bool act = i < a.b.nH;
    const HubChunk h = act ? a.b.hubs[i] : HubChunk{0, 0, 0, 0};
    changed = packed_group<kWave>(a, act, h.v, h.e0, h.e1, tid & 63, true);
  }
  if (__ballot(changed) != 0 && lane_id() == 0) *a.flag = 1;
}

// Sums the chunk counts of every hub into deg_inter and records each chunk's
// offset inside its hub's run of cluster edges.
// ---------------------------------------------------------------------------
// Time-stepped schedule of the same fixed point (RunContractionBFS order with
// the label tie-break): in round t every vertex whose key distance is t --
// final, since only distance t+1 candidates are pushed in round t -- pushes
// key + hop to its neighbours (read first, atomicMin when smaller).  Every
// edge is pushed once per level.  pending[t] is set while some vertex still
// has a distance above t; a round returns at once when the previous round
// found none.
template <int W>
__device__ __forceinline__ bool round_group(const SweepArgs &a, bool active, uint32_t v, uint64_t e0, uint64_t e1,
                                            int gl, uint32_t t, bool hub) {
  const uint64_t kv = active ? a.kin[v] : ~0ull;
  const uint32_t dist = static_cast<uint32_t>(kv >> 32);
  if (active && dist == t) {
    if (!hub) {
      e0 = a.off[v];
      e1 = a.off[v + 1];
    }
    const uint64_t cand = kv + a.hop;
    uint64_t e = e0 + gl;
    for (; e + (kGather - 1) * W < e1; e += kGather * W) {
      uint32_t c[kGather];
      uint64_t k[kGather];
#pragma unroll
      for (int j = 0; j < kGather; ++j) c[j] = a.col[e + j * W];
#pragma unroll
      for (int j = 0; j < kGather; ++j) k[j] = a.kout[c[j]];
#pragma unroll
      for (int j = 0; j < kGather; ++j)
        if (cand < k[j])
          atomicMin(reinterpret_cast<unsigned long long *>(a.kout + c[j]), static_cast<unsigned long long>(cand));
    }
    for (; e < e1; e += W) {
      uint64_t *kw = a.kout + a.col[e];
      if (cand < *kw) atomicMin(reinterpret_cast<unsigned long long *>(kw), static_cast<unsigned long long>(cand));
    }
  }
  return active && dist > t;
}

// Round t's candidates for the largest hubs (degree >= min_deg), pulled
// before the round's pushes: one wave per hub chunk takes the best key + hop
// over the chunk's neighbours at distance t (final in round t) and folds it
// into the hub's key with one atomic.  The pushes that follow read the
// settled key and skip the hub; otherwise thousands of them would each take
// an atomic on that one key, which the device serialises.
__global__ __launch_bounds__(kBlock) void k_hub_pull(SweepArgs a, uint32_t t, uint64_t min_deg) {
  if (a.prev_flag && *a.prev_flag == 0) return;
  const uint32_t i = (blockIdx.x * kBlock + threadIdx.x) / kWave;  // one wave per chunk
  if (i >= a.b.nH) return;
  const HubChunk h = a.b.hubs[i];
  if (a.off[h.v + 1] - a.off[h.v] < min_deg) return;
  const uint64_t kv = a.kin[h.v];
  if (static_cast<uint32_t>(kv >> 32) <= t) return;  // final already
  uint64_t best = ~0ull;
  for (uint64_t e = h.e0 + lane_id(); e < h.e1; e += kWave) {
    const uint64_t ku = a.kin[a.col[e]];
    if (static_cast<uint32_t>(ku >> 32) == t) best = min(best, ku + a.hop);
  }
  for (int o = 32; o > 0; o >>= 1)
    best = min(best, static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(best), o)));
  if (lane_id() == 0 && best < kv)
    atomicMin(reinterpret_cast<unsigned long long *>(a.kout + h.v), static_cast<unsigned long long>(best));
}

// hubs_first: the sections are dispatched in the order hub chunks, wave rows,
// 16-lane rows, one-lane rows (each section's order kept), so that on a
// level stored in index order whose hubs have the smallest indices, as R-MAT's
// do, the pushes with the smallest labels go first.
__global__ __launch_bounds__(kBlock) void k_round(SweepArgs a, uint32_t t, int hubs_first) {
  if (a.prev_flag && *a.prev_flag == 0) return;
  uint32_t blk = blockIdx.x;
  if (hubs_first) {
    const uint32_t nh = a.b.blocks - a.b.bH, nw = a.b.bH - a.b.bW, ns = a.b.bW - a.b.bS;
    blk = blk < nh ? a.b.bH + blk
                   : blk < nh + nw ? a.b.bW + (blk - nh) : blk < nh + nw + ns ? a.b.bS + (blk - nh - nw) : blk - nh - nw - ns;
  }
  const int tid = threadIdx.x;
  bool pending = false;
  if (blk < a.b.bS) {
    const uint32_t i = blk * kBlock + tid;
    const bool act = i < a.b.nT;
    pending = round_group<1>(a, act, act ? a.b.listT[i] : 0, 0, 0, 0, t, false);
  } else if (blk < a.b.bW) {
    const uint32_t i = (blk - a.b.bS) * (kBlock / 16) + tid / 16;
    const bool act = i < a.b.nS;
    pending = round_group<16>(a, act, act ? a.b.listS[i] : 0, 0, 0, tid & 15, t, false);
  } else if (blk < a.b.bH) {
    const uint32_t i = (blk - a.b.bW) * (kBlock / kWave) + tid / kWave;
    const bool act = i < a.b.nW;
    pending = round_group<kWave>(a, act, act ? a.b.listW[i] : 0, 0, 0, tid & 63, t, false);
  } else {
    const uint32_t i = (blk - a.b.bH) * (kBlock / kWave) + tid / kWave;
    const bool act = i < a.b.nH;
    HubChunk h = act ? a.b.hubs[i] : HubChunk{0, 0, 0, 0};
    pending = round_group<kWave>(a, act, h.v, h.e0, h.e1, tid & 63, t, true);
  }
  if (__ballot(pending) != 0 && lane_id() == 0) *a.flag = 1;
}

// Relabel pass: the labels of col[e] that differ from the row's label are
// written packed at the start of the row's slots (a hub chunk's at the start
// of the chunk), in ballot order; the count goes to deg_inter (hub chunks:
// hub_inter).  The slots after the count are left as they were.
// Groups of 16 or 64 lanes also drop repeats of a label inside their row
// through a small direct-mapped LDS filter whose slots are claimed by
// compare-and-swap (best effort: a label is dropped only when the filter
// already holds it or another lane's swap put it there in the same step, so
// every kept set still covers the row's distinct labels; the bucket dedup
// removes the rest).
// Label lookups: the first n_hot entries (the highest-degree vertices of a
// degree-ordered level, which receive most edge endpoints) from an LDS copy,
// the others from the array.
struct LabHot {
  const uint32_t *__restrict__ lab;
  const uint32_t *hot;
  uint32_t n_hot;
  __device__ __forceinline__ uint32_t operator()(uint32_t c) const { return c < n_hot ? hot[c] : lab[c]; }
  __device__ __forceinline__ uint32_t row(uint32_t v) const { return (*this)(v); }
};

// The same with 16-bit entries: the first n_hot (up to 2 kRelabelHot)
// positions of a degree-ordered level 0, whose clusters were numbered first
// (cluster_ranks_hot), so that their ids fit 16 bits.
struct LabHot16 {
  const uint32_t *__restrict__ lab;
  const uint16_t *hot;
  uint32_t n_hot;
  __device__ __forceinline__ uint32_t operator()(uint32_t c) const { return c < n_hot ? hot[c] : lab[c]; }
  __device__ __forceinline__ uint32_t row(uint32_t v) const { return (*this)(v); }
};

// Label lookups of one owner part p (the vertices t with owner_of(t) == p, a
// 128-byte line of labels in every kParts lines): the part's share of the
// first kSplitHot vertices from LDS, packed line by line, the others from the
// array.  Only targets of part p are looked up through it; the row's own
// label (any part) comes from the array.
constexpr uint32_t kRelabelHot = 28672;
constexpr uint32_t kRelabelHot16 = 2 * kRelabelHot;  // 16-bit entries in the same LDS
static_assert(kRelabelHot16 <= kHotClusters, "16-bit LDS labels need the hot positions' clusters numbered first");
constexpr uint32_t kSplitHot = kParts * kRelabelHot;
__device__ __forceinline__ uint32_t part_slot(uint32_t t) { return ((t >> 8) << 5) | (t & 31); }
struct LabPart {
  const uint32_t *__restrict__ lab;
  const uint32_t *hot;
  __device__ __forceinline__ uint32_t operator()(uint32_t c) const { return c < kSplitHot ? hot[part_slot(c)] : lab[c]; }
  __device__ __forceinline__ uint32_t row(uint32_t v) const { return lab[v]; }
};

template <int W, class Lab>
__device__ __forceinline__ void relabel_group(const SweepArgs &a, const Lab &lab,
                                              uint32_t *relab, bool active, uint32_t v,
                                              uint64_t e0, uint64_t e1, int gl, bool hub, uint32_t chunk,
                                              uint32_t *filt, uint32_t fmask) {
  if (W > 1) {
    for (uint32_t j = gl; j <= fmask; j += W) filt[j] = kNone;
    wave_fence();
  }
  const int gbase = static_cast<int>(lane_id()) & ~(W - 1);
  const uint64_t gmask = (W >= 64) ? ~0ull : (((1ull << (W & 63)) - 1ull) << gbase);
  const uint32_t lv = active ? lab.row(v) : 0u;
  // every lane of the wave takes the same number of W-edge steps (the most any
  // group needs), so the ballots below see every lane of each group
  uint64_t rounds = active ? ceil_div(e1 - e0, static_cast<uint64_t>(W)) : 0;
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t t = static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(rounds), o, 64));
    rounds = t > rounds ? t : rounds;
  }
  uint32_t base = 0;  // cluster edges of the row (chunk) written so far
  for (uint64_t r = 0; r < rounds; r += kRelabelGather) {
    uint32_t c[kRelabelGather], l[kRelabelGather];
    bool ok[kRelabelGather];
#pragma unroll
    for (int j = 0; j < kRelabelGather; ++j) {
      const uint64_t e = e0 + (r + j) * W + gl;
      ok[j] = active && e < e1;
      c[j] = ok[j] ? a.col[e] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kRelabelGather; ++j) l[j] = ok[j] ? lab(c[j]) : lv;
#pragma unroll
    for (int j = 0; j < kRelabelGather; ++j) {
      bool inter = l[j] != lv;
      if (W > 1 && inter) {
        // the slot is claimed by compare-and-swap, so of the group's lanes
        // that meet one label in this step only the one whose swap lands
        // keeps it (a lane whose swap meets another label keeps its own)
        const uint32_t slot = hash32(l[j]) & fmask;
        const uint32_t cur = filt[slot];
        if (cur == l[j] || atomicCAS(&filt[slot], cur, l[j]) == l[j]) inter = false;
      }
      const uint64_t m = __ballot(inter) & gmask;
      if (inter) relab[e0 + base + __popcll(m & lanes_below())] = l[j];
      base += static_cast<uint32_t>(__popcll(m));
    }
  }
  if (active && gl == 0) {
    if (hub)
      a.hub_inter[chunk] = base;
    else
      a.deg_inter[v] = base;
  }
}

constexpr int kFilterWords = 2048;  // 8 KB of LDS per block

// One 256-thread block's share of the relabel pass: section by block index.
template <class Lab>
__device__ __forceinline__ void relabel_block(const SweepArgs &a, const Lab &lab, uint32_t *relab, uint32_t blk,
                                              int tid, uint32_t *filt) {
  if (blk < a.b.bS) {
    const uint32_t i = blk * kBlock + tid;
    const bool act = i < a.b.nT;
    uint32_t v = act ? a.b.listT[i] : 0;
    uint64_t e0 = act ? a.off[v] : 0, e1 = act ? a.off[v + 1] : 0;
    relabel_group<1>(a, lab, relab, act, v, e0, e1, 0, false, 0, nullptr, 0);
  } else if (blk < a.b.bW) {
    const uint32_t i = (blk - a.b.bS) * (kBlock / 16) + tid / 16;
    const bool act = i < a.b.nS;
    uint32_t v = act ? a.b.listS[i] : 0;
    uint64_t e0 = act ? a.off[v] : 0, e1 = act ? a.off[v + 1] : 0;
    relabel_group<16>(a, lab, relab, act, v, e0, e1, tid & 15, false, 0, filt + (tid / 16) * 128, 127);
  } else if (blk < a.b.bH) {
    const uint32_t i = (blk - a.b.bW) * (kBlock / kWave) + tid / kWave;
    const bool act = i < a.b.nW;
    uint32_t v = act ? a.b.listW[i] : 0;
    uint64_t e0 = act ? a.off[v] : 0, e1 = act ? a.off[v + 1] : 0;
    relabel_group<kWave>(a, lab, relab, act, v, e0, e1, tid & 63, false, 0, filt + (tid / kWave) * 512, 511);
  } else {
    const uint32_t i = (blk - a.b.bH) * (kBlock / kWave) + tid / kWave;
    const bool act = i < a.b.nH;
    HubChunk h = act ? a.b.hubs[i] : HubChunk{0, 0, 0, 0};
    relabel_group<kWave>(a, lab, relab, act, h.v, h.e0, h.e1, tid & 63, true, i, filt + (tid / kWave) * 512, 511);
  }
}

// Persistent launch: each 1024-thread block first copies the labels of the
// first n_hot vertices into LDS (the highest-degree ones on a degree-ordered
// level; n_hot = 0 elsewhere), then its four 256-thread quarters walk the
// virtual 256-thread blocks of the binned layout in a grid stride, each
// quarter with its own LDS filter (the quarters are whole waves and need no
// barrier between them).
constexpr int kRelabelQuarters = 4;

// H16: 16-bit entries for twice the positions (their cluster ids fit 16 bits).
template <bool H16>
__global__ __launch_bounds__(kBlock * kRelabelQuarters) void k_relabel_hot(SweepArgs a, const uint32_t *lab,
                                                                        uint32_t *relab, uint32_t n_hot) {
  __shared__ uint32_t hot[kRelabelHot];
  __shared__ uint32_t filt[kRelabelQuarters][kFilterWords];
  const int q = threadIdx.x / kBlock, tid = threadIdx.x % kBlock;
  if (H16) {
    uint16_t *h16 = reinterpret_cast<uint16_t *>(hot);
    for (uint32_t i = threadIdx.x; i < n_hot; i += kBlock * kRelabelQuarters) h16[i] = static_cast<uint16_t>(lab[i]);
    __syncthreads();
    const LabHot16 lk{lab, h16, n_hot};
    for (uint32_t blk = blockIdx.x * kRelabelQuarters + q; blk < a.b.blocks; blk += gridDim.x * kRelabelQuarters)
      relabel_block(a, lk, relab, blk, tid, filt[q]);
  } else {
    for (uint32_t i = threadIdx.x; i < n_hot; i += kBlock * kRelabelQuarters) hot[i] = lab[i];
    __syncthreads();
    const LabHot lk{lab, hot, n_hot};
    for (uint32_t blk = blockIdx.x * kRelabelQuarters + q; blk < a.b.blocks; blk += gridDim.x * kRelabelQuarters)
      relabel_block(a, lk, relab, blk, tid, filt[q]);
  }
}

// Owner-split relabel of a degree-ordered level 0.  The rows above the 16-lane
// bin are cut into pieces, one per owner group of the row (cut again every
// wmax slots), which the launch spreads over kParts parts of its blocks: the
// blocks of part p (blockIdx % kParts == p; the device hands consecutive
// workgroups to its kParts chiplets in turn, each with its own L2) look up
// only labels of part p, so each chiplet's L2 serves one part's labels and its
// LDS holds that part's share of the kSplitHot highest-degree vertices.  Every
// block first runs its share of the one-lane and 16-lane bins (rows kept
// whole) with the first n_hot labels in LDS, as k_relabel_hot, then reloads
// LDS with its part's share and runs its part's pieces, one wave per piece.
// Results do not depend on which chiplet runs a block.  The first n_hot
// positions' cluster ids must fit 16 bits (cluster_ranks_hot).
struct PartRange {
  // pieces of part p: part_list[at[p] .. at[p + 1]); p = kParts: the shared
  // pieces (whole rows up to the split bound), run by every block
  uint32_t at[kParts + 2];
};

__global__ __launch_bounds__(kBlock * kRelabelQuarters) void k_relabel_split(SweepArgs a, const uint32_t *lab,
                                                                          uint32_t *relab, uint32_t n_hot, uint64_t n,
                                                                          const HubChunk *pieces,
                                                                          const uint32_t *part_list, PartRange pr) {
  __shared__ uint32_t hot[kRelabelHot];
  __shared__ uint32_t filt[kRelabelQuarters][kFilterWords];
  const int q = threadIdx.x / kBlock, tid = threadIdx.x % kBlock;
  uint16_t *h16 = reinterpret_cast<uint16_t *>(hot);
  for (uint32_t i = threadIdx.x; i < n_hot; i += kBlock * kRelabelQuarters) h16[i] = static_cast<uint16_t>(lab[i]);
  __syncthreads();
  constexpr uint32_t kWavesPerBlock = kBlock * kRelabelQuarters / kWave;
  const uint32_t wave = threadIdx.x / kWave;
  uint32_t *wfilt = filt[q] + (tid / kWave) * 512;
  {
    const LabHot16 lk{lab, h16, n_hot};
    for (uint32_t blk = blockIdx.x * kRelabelQuarters + q; blk < a.b.bW; blk += gridDim.x * kRelabelQuarters)
      relabel_block(a, lk, relab, blk, tid, filt[q]);
    const uint32_t s0 = pr.at[kParts], shared = pr.at[kParts + 1] - s0;
    for (uint32_t i = blockIdx.x * kWavesPerBlock + wave; i < shared; i += gridDim.x * kWavesPerBlock) {
      const uint32_t c = part_list[s0 + i];
      const HubChunk h = pieces[c];
      relabel_group<kWave>(a, lk, relab, true, h.v, h.e0, h.e1, tid & 63, true, c, wfilt, 511);
    }
  }
  __syncthreads();
  const uint32_t part = blockIdx.x % kParts;
  for (uint32_t i = threadIdx.x; i < kRelabelHot; i += kBlock * kRelabelQuarters) {
    const uint64_t t = (static_cast<uint64_t>(i >> 5) << 8) | (part << 5) | (i & 31);
    hot[i] = t < n ? lab[t] : kNone;
  }
  __syncthreads();
  const LabPart lp{lab, hot};
  const uint32_t p0 = pr.at[part], count = pr.at[part + 1] - p0;
  const uint32_t stride = (gridDim.x / kParts) * kWavesPerBlock;
  for (uint32_t i = (blockIdx.x / kParts) * kWavesPerBlock + wave; i < count; i += stride) {
    const uint32_t c = part_list[p0 + i];
    const HubChunk h = pieces[c];
    relabel_group<kWave>(a, lp, relab, true, h.v, h.e0, h.e1, tid & 63, true, c, wfilt, 511);
  }
}

// Pieces of the rows above smax: bounds of owner group x inside a row whose
// targets are grouped by owner (binary search for the first slot of part >= x).
__device__ __forceinline__ uint64_t owner_bound(const uint32_t *col, uint64_t lo, uint64_t hi, uint32_t x) {
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (owner_of(col[mid]) < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// cnt[v] = pieces of row v: 0 up to smax slots; up to split slots the whole
// row cut every wmax slots (shared pieces); above, the sum over its owner
// groups of ceil(group slots / wmax).
__global__ __launch_bounds__(kBlock) void k_piece_count(uint64_t n, const uint64_t *off, const uint32_t *col,
                                                        uint32_t smax, uint32_t wmax, uint64_t split, uint32_t *cnt) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const uint64_t e0 = off[v], e1 = off[v + 1];
    uint64_t c = 0;
    if (e1 - e0 > smax && e1 - e0 <= split) {
      c = ceil_div(e1 - e0, wmax);
    } else if (e1 - e0 > smax) {
      uint64_t lo = e0;
      for (uint32_t x = 1; x <= kParts; ++x) {
        const uint64_t hi = x < kParts ? owner_bound(col, lo, e1, x) : e1;
        c += ceil_div(hi - lo, wmax);
        lo = hi;
      }
    }
    cnt[v] = static_cast<uint32_t>(c);
  }
}

// The pieces of every row from start[v] (exclusive prefix of the counts), in
// row order and within a row by owner group, then slot: the chunks of one row
// are consecutive with k = 0, 1, ... as hub chunks are.
__global__ __launch_bounds__(kBlock) void k_piece_write(uint64_t n, const uint64_t *off, const uint32_t *col,
                                                        uint32_t smax, uint32_t wmax, uint64_t split,
                                                        const uint64_t *start, HubChunk *pieces, uint8_t *part) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const uint64_t e0 = off[v], e1 = off[v + 1];
    if (e1 - e0 <= smax) continue;
    uint64_t p = start[v];
    uint32_t k = 0;
    if (e1 - e0 <= split) {
      for (uint64_t e = e0; e < e1; e += wmax) {
        pieces[p] = HubChunk{static_cast<uint32_t>(v), k, e, e + wmax < e1 ? e + wmax : e1};
        part[p] = static_cast<uint8_t>(kParts);
        ++p;
        ++k;
      }
      continue;
    }
    uint64_t lo = e0;
    for (uint32_t x = 1; x <= kParts; ++x) {
      const uint64_t hi = x < kParts ? owner_bound(col, lo, e1, x) : e1;
      for (uint64_t e = lo; e < hi; e += wmax) {
        pieces[p] = HubChunk{static_cast<uint32_t>(v), k, e, e + wmax < hi ? e + wmax : hi};
        part[p] = static_cast<uint8_t>(x - 1);
        ++p;
        ++k;
      }
      lo = hi;
    }
  }
}

// Owner-grouped sort keys over B >= 8 bits (t < 2^B): the owner in the top
// three bits, then the target without its owner bits, so that ascending keys
// list a row's targets by owner group, ascending inside a group.
__device__ __forceinline__ uint32_t owner_key(uint32_t t, int B) {
  return (owner_of(t) << (B - 3)) | ((t >> 8) << 5) | (t & 31);
}
__device__ __forceinline__ uint32_t owner_unkey(uint32_t k, int B) {
  const uint32_t x = k >> (B - 3), r = k & ((1u << (B - 3)) - 1u);
  return ((r >> 5) << 8) | (x << 5) | (r & 31);
}
__global__ __launch_bounds__(kBlock) void k_owner_keys(uint64_t m2, const uint32_t *col, uint32_t *key, int B) {
  for (uint64_t e = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; e < m2;
       e += static_cast<uint64_t>(gridDim.x) * kBlock)
    key[e] = owner_key(col[e], B);
}
__global__ __launch_bounds__(kBlock) void k_owner_cols(uint64_t m2, const uint32_t *key, uint32_t *col, int B) {
  for (uint64_t e = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; e < m2;
       e += static_cast<uint64_t>(gridDim.x) * kBlock)
    col[e] = owner_unkey(key[e], B);
}

// From the exclusive prefix of all chunk counts: a chunk's offset in its hub's
// run is its prefix minus the hub's first chunk's, and the hub's last chunk
// writes the hub's total (chunks of one hub are consecutive).
__global__ void k_hub_fold(const HubChunk *hubs, uint32_t nH, const uint64_t *prefix, uint32_t *deg_inter,
                           uint32_t *hub_pos) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nH) return;
  const HubChunk h = hubs[c];
  const uint64_t first = prefix[c - h.k];
  hub_pos[c] = static_cast<uint32_t>(prefix[c] - first);
  if (c + 1 == nH || hubs[c + 1].k == 0) deg_inter[h.v] = static_cast<uint32_t>(prefix[c + 1] - first);
}

__global__ void k_jump(uint64_t n, uint64_t *key, int *flag) {
  uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  bool changed = false;
  if (v < n) {
    const uint64_t k = key[v];
    const uint64_t k2 = key[static_cast<uint32_t>(k)];
    if (k2 < k) {
      key[v] = k2;
      changed = true;
    }
  }
  if (__ballot(changed) != 0 && lane_id() == 0) *flag = 1;
}

// ---------------------------------------------------------------------------
// Blocks of consecutive level indices (block b holds [b*bsize, (b+1)*bsize)),
// the PEs of a vertex-block distribution; their components come from
// kSweepBlock sweeps with pointer jumping.
// Interface vertices (a neighbour outside the block) keep their own index as
// label: key[v] = v.
__global__ __launch_bounds__(kBlock) void k_block_interface(const uint64_t *off, const uint32_t *col, uint64_t n,
                                                           uint64_t bsize, uint64_t *key) {
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / kWave);
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kBlock / kWave) + threadIdx.x / kWave; v < n; v += waves) {
    const uint64_t lo = v / bsize * bsize, hi = lo + bsize < n ? lo + bsize : n;
    bool out = false;
    const uint64_t e1 = off[v + 1];
    for (uint64_t e = off[v] + lane_id(); e < e1; e += kWave) {
      const uint32_t u = col[e];
      out = out || u < lo || u >= hi;
    }
    if (__ballot(out) != 0 && lane_id() == 0) key[v] = v;
  }
}

// ---------------------------------------------------------------------------
// Centre labels and the maximum cluster radius (hops from the centre).
// lab[v] = the centre's level index; the centre's key is at pos[l] when the
// level is stored in another order (null: l).
// Grid-stride, with one atomic per workgroup: atomics on one address are
// serialised by the device, so one per wave cost more than the pass itself.
__global__ __launch_bounds__(kBlock) void k_labels_radius(uint64_t n, const uint64_t *key, const uint32_t *pos,
                                                         uint32_t *lab, uint32_t *used, uint32_t *radius,
                                                         int with_radius) {
  __shared__ uint32_t wave_max[kBlock / kWave];
  uint32_t r = 0;
  // four vertices per lane per step, so that their chains of dependent loads
  // (key, centre position, centre key) are in flight together
  constexpr int kU = 4;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
  for (uint64_t v0 = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; v0 < n; v0 += kU * stride) {
    uint64_t k[kU];
    uint32_t c[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) k[u] = v0 + u * stride < n ? key[v0 + u * stride] : 0ull;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const uint64_t v = v0 + u * stride;
      if (v >= n) continue;
      const uint32_t l = static_cast<uint32_t>(k[u]);
      lab[v] = l;
      if (used) used[l] = 1;
    }
    if (with_radius) {
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const uint32_t l = static_cast<uint32_t>(k[u]);
        c[u] = v0 + u * stride < n ? (pos ? pos[l] : l) : 0u;
      }
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (v0 + u * stride < n)
          r = max(r, static_cast<uint32_t>(k[u] >> 32) - static_cast<uint32_t>(key[c[u]] >> 32));
    }
  }
  if (with_radius) {
    for (int o = 32; o > 0; o >>= 1) r = max(r, static_cast<uint32_t>(__shfl_xor(r, o)));
    if (lane_id() == 0) wave_max[threadIdx.x / kWave] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < kBlock / kWave; ++w) r = max(r, wave_max[w]);
      if (r) atomicMax(radius, r);
    }
  }
}

// ---------------------------------------------------------------------------
// Generic three-phase exclusive scan over functor-defined values.
constexpr int kScanPerThread = 8;
constexpr int kScanTile = kBlock * kScanPerThread;
constexpr uint64_t kScanFusedBlocks = 1024;  // two-launch scans up to this many tiles

struct Pair64 {
  uint64_t a, b;
  __host__ __device__ Pair64 operator+(const Pair64 &o) const { return {a + o.a, b + o.b}; }
};

// Block-wide exclusive scan (Hillis-Steele in LDS); *total receives the block sum.
template <class T>
__device__ T block_exclusive_scan_shift(T v, T *smem, T *total) {
  const int tid = threadIdx.x;
  smem[tid] = v;
  __syncthreads();
  for (int o = 1; o < kBlock; o <<= 1) {
    T t = tid >= o ? smem[tid - o] : T{};
    __syncthreads();
    if (tid >= o) smem[tid] = smem[tid] + t;
    __syncthreads();
  }
  T excl = tid > 0 ? smem[tid - 1] : T{};
  T tot = smem[kBlock - 1];
  __syncthreads();
  if (total) *total = tot;
  return excl;
}

template <class T, class Load>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(uint64_t n, Load load, T *block_sums) {
  __shared__ T smem[kBlock];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  T acc{};
  for (int j = 0; j < kScanPerThread; ++j) {
    uint64_t i = base + static_cast<uint64_t>(j) * kBlock + threadIdx.x;
    if (i < n) acc = acc + load(i);
  }
  T tot;
  block_exclusive_scan_shift<T>(acc, smem, &tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__device__ __forceinline__ uint64_t shfl_up_t(uint64_t x, int o) {
  return static_cast<uint64_t>(__shfl_up(static_cast<unsigned long long>(x), o));
}
__device__ __forceinline__ Pair64 shfl_up_t(Pair64 x, int o) { return {shfl_up_t(x.a, o), shfl_up_t(x.b, o)}; }

// Sum of one value per thread over the block (wave scans, then the wave sums
// through LDS); every thread receives it.
template <class T>
__device__ T block_total(T v, T *wsum) {
  const uint32_t ln = lane_id(), wv = threadIdx.x / kWave;
  for (int o = 1; o < kWave; o <<= 1) {
    const T t = shfl_up_t(v, o);
    if (ln >= static_cast<uint32_t>(o)) v = v + t;
  }
  if (ln == kWave - 1) wsum[wv] = v;
  __syncthreads();
  T tot{};
  for (uint32_t w = 0; w < kBlock / kWave; ++w) tot = tot + wsum[w];
  __syncthreads();
  return tot;
}

// A tile is kScanPerThread rounds of kBlock consecutive items (lane i of a
// round takes item i, so loads and stores are coalesced); each round is one
// block-wide scan (wave scans, then the wave sums through LDS), carried into
// the next.  block_prefix: the exclusive prefix of every block's tile; or
// block_sums (block_prefix null): every block's tile sum, the block adding up
// those before its own (for grids of up to kScanFusedBlocks blocks, in place
// of a scan launch over the sums).
template <class T, class Load, class Store>
__global__ __launch_bounds__(kBlock) void k_scan_down(uint64_t n, Load load, Store store, const T *block_prefix,
                                                     T *total_out, const T *block_sums = nullptr) {
  __shared__ T wsum[kBlock / kWave];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  const uint32_t ln = lane_id(), wv = threadIdx.x / kWave;
  T run{};
  if (block_prefix) {
    run = block_prefix[blockIdx.x];
  } else if (block_sums) {
    T acc{};
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kBlock) acc = acc + block_sums[b];
    run = block_total<T>(acc, wsum);
  }
  for (int j = 0; j < kScanPerThread; ++j) {
    const uint64_t i = base + static_cast<uint64_t>(j) * kBlock + threadIdx.x;
    const T v = i < n ? load(i) : T{};
    T inc = v;
    for (int o = 1; o < kWave; o <<= 1) {
      const T t = shfl_up_t(inc, o);
      if (ln >= static_cast<uint32_t>(o)) inc = inc + t;
    }
    T ex = shfl_up_t(inc, 1);
    if (ln == 0) ex = T{};
    if (ln == kWave - 1) wsum[wv] = inc;
    __syncthreads();
    T before{}, round_total{};
    for (uint32_t w = 0; w < kBlock / kWave; ++w) {
      const T x = wsum[w];
      if (w < wv) before = before + x;
      round_total = round_total + x;
    }
    __syncthreads();
    if (i < n) store(i, run + before + ex, v);
    run = run + round_total;
  }
  if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total_out = run;
}

template <class T>
struct ArrayLoad {
  const T *p;
  __device__ T operator()(uint64_t i) const { return p[i]; }
};
struct ArrLoadU32 {
  const uint32_t *p;
  __device__ uint64_t operator()(uint64_t i) const { return p[i]; }
};
template <class T>
struct ArrayStore {
  T *p;
  __device__ void operator()(uint64_t i, T prefix, T) const { p[i] = prefix; }
};

// Scratch words (u64) needed by scan_exclusive for n elements.
uint64_t scan_words_for(uint64_t n, size_t elem_words) {
  uint64_t words = 0;
  uint64_t blocks = ceil_div(n, kScanTile);
  while (blocks > 1) {
    words += 2 * blocks * elem_words + 8;
    blocks = ceil_div(blocks, kScanTile);
  }
  return words + 16 * elem_words;
}

// Exclusive scan; the total lands in *total (device).  scratch must hold
// scan_words_for(n) u64 words.
template <class T, class Load, class Store>
void scan_exclusive(uint64_t n, Load load, Store store, T *total, uint64_t *scratch, hipStream_t s) {
  if (n == 0) {
    hipLaunchKernelGGL((k_fill<T>), dim3(1), dim3(1), 0, s, total, 1, T{});
    KC_LAUNCH_CHECK();
    return;
  }
  const uint64_t blocks = ceil_div(n, kScanTile);
  if (blocks == 1) {
    hipLaunchKernelGGL((k_scan_down<T, Load, Store>), dim3(1), dim3(kBlock), 0, s, n, load, store,
                       static_cast<const T *>(nullptr), total);
    KC_LAUNCH_CHECK();
    return;
  }
  T *sums = reinterpret_cast<T *>(scratch);
  T *prefix = sums + blocks;
  uint64_t *rest = reinterpret_cast<uint64_t *>(prefix + blocks) + 4;
  hipLaunchKernelGGL((k_scan_reduce<T, Load>), dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, n, load,
                     sums);
  KC_LAUNCH_CHECK();
  if (blocks <= kScanFusedBlocks) {
    // each block adds up the tile sums before its own: two launches, not three
    hipLaunchKernelGGL((k_scan_down<T, Load, Store>), dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, n,
                       load, store, static_cast<const T *>(nullptr), total, static_cast<const T *>(sums));
    KC_LAUNCH_CHECK();
    return;
  }
  scan_exclusive<T>(blocks, ArrayLoad<T>{sums}, ArrayStore<T>{prefix}, total, rest, s);
  hipLaunchKernelGGL((k_scan_down<T, Load, Store>), dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, n,
                     load, store, static_cast<const T *>(prefix), static_cast<T *>(nullptr));
  KC_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Clusters from the (label, vertex) list sorted by label: every vertex with a
// cluster edge is selected, the list is sorted, and one scan over the sorted
// list yields each cluster's next-level index (segment number), its bucket
// offset and every vertex's write position -- no atomics.
struct SelectLoad {
  const uint32_t *deg_inter;
  __device__ uint64_t operator()(uint64_t i) const { return deg_inter[i] != 0 ? 1ull : 0ull; }
};
struct SelectStore {
  const uint64_t *key;
  uint32_t *sk, *sv;
  __device__ void operator()(uint64_t i, uint64_t prefix, uint64_t v) const {
    if (v) {
      sk[prefix] = static_cast<uint32_t>(key[i]);
      sv[prefix] = static_cast<uint32_t>(i);
    }
  }
};

// (segment head, cluster-edge count) of item i.  The first pass of the scan
// gathers the count by vertex and keeps it in `cache` in item order; the
// second pass reads it from there instead of gathering again.
struct SegLoad {
  const uint32_t *sk, *sv, *deg_inter;
  uint32_t *cache;
  __device__ Pair64 operator()(uint64_t i) const {
    const uint32_t a = sk[i];
    const bool head = i == 0 || sk[i - 1] != a;
    const uint32_t d = deg_inter[sv[i]];
    cache[i] = d;
    return {head ? 1ull : 0ull, d};
  }
};
struct SegLoadCached {
  const uint32_t *sk, *cache;
  __device__ Pair64 operator()(uint64_t i) const {
    const uint32_t a = sk[i];
    const bool head = i == 0 || sk[i - 1] != a;
    return {head ? 1ull : 0ull, cache[i]};
  }
};
struct SegStore {
  const uint32_t *sk, *sv, *gid, *rank;
  uint32_t *pos_of, *newidx, *seg_of_cid, *gid_next;
  uint64_t *seg_off;
  __device__ void operator()(uint64_t i, Pair64 prefix, Pair64 v) const {
    pos_of[sv[i]] = static_cast<uint32_t>(prefix.b);
    if (v.a) {
      const uint32_t a = sk[i];
      const uint32_t seg = static_cast<uint32_t>(prefix.a);
      newidx[a] = seg;
      seg_of_cid[rank[a]] = seg;
      gid_next[seg] = gid ? gid[a] : a;
      seg_off[seg] = prefix.b;
    }
  }
};

__global__ void k_seg_finish(const Pair64 *total, uint64_t *seg_off, uint64_t *sizes) {
  seg_off[total->a] = total->b;
  sizes[0] = total->a;
  sizes[1] = total->b;
}
__global__ void k_copy_total(const uint64_t *t, uint64_t *dst) { *dst = *t; }

// ---------------------------------------------------------------------------
// Scatter of cluster edges into buckets (binned like the sweep).
// Gather of every row's cluster edges (the non-kNone relabel slots) into the
// bucket of its cluster, at the row's position: sequential reads of the row,
// contiguous writes.
struct ScatterArgs {
  const uint64_t *off;
  const uint32_t *relab;
  Bins b;
  const uint32_t *deg_inter;
  const uint32_t *hub_inter;
  const uint32_t *hub_pos;
  const uint32_t *pos_of;
  uint32_t *bucket;
};

template <int W>
__device__ __forceinline__ void scatter_group(const ScatterArgs &a, bool active, uint32_t v, uint64_t e0,
                                              uint64_t e1, int gl, uint32_t d, uint32_t extra) {
  (void)e1;
  // group base lane within the wave
  const int gbase = static_cast<int>(lane_id()) & ~(W - 1);
  uint64_t pos = 0;
  active = active && d > 0;
  if (active && gl == 0) pos = static_cast<uint64_t>(a.pos_of[v]) + extra;
  if (W > 1) pos = static_cast<uint64_t>(__shfl(static_cast<unsigned long long>(pos), gbase, 64));
  if (!active) return;
  // the row's (chunk's) d cluster edges sit at its start: a straight copy,
  // four loads in flight per lane before their stores
  uint32_t i = gl;
  for (; i + 3 * W < d; i += 4 * W) {
    uint32_t x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) x[j] = a.relab[e0 + i + j * W];
#pragma unroll
    for (int j = 0; j < 4; ++j) a.bucket[pos + i + j * W] = x[j];
  }
  for (; i < d; i += W) a.bucket[pos + i] = a.relab[e0 + i];
}

// Rows of the one-lane bin (16 slots or fewer with the default bins): each
// lane reads one row's bounds, then the wave copies its 64 rows in 16 steps
// of four rows, 16 lanes to a row, all loads issued before the stores, so
// that the copies are contiguous per row instead of one slot per lane.  A
// wave holding a longer row (larger one-lane bins) copies lane by lane.
__device__ __forceinline__ void scatter_small_rows(const ScatterArgs &a, bool active, uint32_t v) {
  constexpr int kSteps = kWave / 4;
  const uint32_t ln = lane_id(), grp = ln / 16, gl = ln % 16;
  const uint64_t e0 = active ? a.off[v] : 0;
  const uint32_t d = active ? a.deg_inter[v] : 0;
  if (__ballot(d > 16) != 0) {
    scatter_group<1>(a, active, v, e0, 0, 0, d, 0);
    return;
  }
  const uint64_t pos = d ? static_cast<uint64_t>(a.pos_of[v]) : 0;
  uint32_t val[kSteps];
#pragma unroll
  for (int p = 0; p < kSteps; ++p) {
    const int src = p * 4 + static_cast<int>(grp);
    const uint32_t dp = static_cast<uint32_t>(__shfl(static_cast<int>(d), src));
    const uint64_t ep = static_cast<uint64_t>(__shfl(static_cast<unsigned long long>(e0), src));
    val[p] = gl < dp ? a.relab[ep + gl] : 0u;
  }
#pragma unroll
  for (int p = 0; p < kSteps; ++p) {
    const int src = p * 4 + static_cast<int>(grp);
    const uint32_t dp = static_cast<uint32_t>(__shfl(static_cast<int>(d), src));
    const uint64_t pp = static_cast<uint64_t>(__shfl(static_cast<unsigned long long>(pos), src));
    if (gl < dp) a.bucket[pp + gl] = val[p];
  }
}

__global__ __launch_bounds__(kBlock) void k_scatter(ScatterArgs a) {
  const uint32_t blk = blockIdx.x;
  const int tid = threadIdx.x;
  if (blk < a.b.bS) {
    const uint32_t i = blk * kBlock + tid;
    const bool act = i < a.b.nT;
    scatter_small_rows(a, act, act ? a.b.listT[i] : 0);
  } else if (blk < a.b.bW) {
    const uint32_t i = (blk - a.b.bS) * (kBlock / 16) + tid / 16;
    const bool act = i < a.b.nS;
    uint32_t v = act ? a.b.listS[i] : 0;
    uint64_t e0 = act ? a.off[v] : 0, e1 = act ? a.off[v + 1] : 0;
    scatter_group<16>(a, act, v, e0, e1, tid & 15, act ? a.deg_inter[v] : 0, 0);
  } else if (blk < a.b.bH) {
    const uint32_t i = (blk - a.b.bW) * (kBlock / kWave) + tid / kWave;
    const bool act = i < a.b.nW;
    uint32_t v = act ? a.b.listW[i] : 0;
    uint64_t e0 = act ? a.off[v] : 0, e1 = act ? a.off[v + 1] : 0;
    scatter_group<kWave>(a, act, v, e0, e1, tid & 63, act ? a.deg_inter[v] : 0, 0);
  } else {
    const uint32_t i = (blk - a.b.bH) * (kBlock / kWave) + tid / kWave;
    const bool act = i < a.b.nH;
    HubChunk h = act ? a.b.hubs[i] : HubChunk{0, 0, 0, 0};
    scatter_group<kWave>(a, act, h.v, h.e0, h.e1, tid & 63, act ? a.hub_inter[i] : 0, act ? a.hub_pos[i] : 0);
  }
}

// ---------------------------------------------------------------------------
// Bucket dedup.  Small buckets: one wave, open-addressing table in LDS.
constexpr int kWaveTab = 2 * kDedupWaveCap;   // 1024 slots per wave
constexpr int kBlockTab = 2 * kDedupBlockCap;  // 16384 slots per block (64 KB)

// Open addressing with linear probing; a slot is read before the CAS so that
// duplicates (frequent in hub buckets) cost a load instead of an atomic.
__device__ __forceinline__ bool lds_insert(uint32_t *tab, uint32_t mask, uint32_t b) {
  uint32_t h = hash32(b) & mask;
  while (true) {
    const uint32_t cur = tab[h];
    if (cur == b) return false;
    if (cur == kNone) {
      const uint32_t prev = atomicCAS(&tab[h], kNone, b);
      if (prev == kNone) return true;
      if (prev == b) return false;
    }
    h = (h + 1) & mask;
  }
}

// Medium and huge buckets listed by a scan (positions from the prefix, no
// atomics: appends through one shared counter are serialised by the device).
struct ClassLoad {
  const uint64_t *n_dev;
  uint64_t n;
  const uint64_t *seg_off;
  __device__ Pair64 operator()(uint64_t i) const {
    if (i >= (n_dev ? *n_dev : n)) return Pair64{0, 0};
    const uint64_t s = seg_off[i + 1] - seg_off[i];
    return s <= kDedupWaveCap ? Pair64{0, 0} : s <= kDedupBlockCap ? Pair64{1, 0} : Pair64{0, 1};
  }
};
struct ClassStore {
  uint32_t *medium, *huge;
  __device__ void operator()(uint64_t i, Pair64 prefix, Pair64 v) const {
    if (v.a) medium[prefix.a] = static_cast<uint32_t>(i);
    if (v.b) huge[prefix.b] = static_cast<uint32_t>(i);
  }
};

// Small buckets, four consecutive ones per wave step.  Those of at most 16
// elements (most buckets) take 16 lanes each, one element per lane, and a
// lane keeps its element when no lower lane of its group holds the same
// value: no table, no LDS.  The wave then takes the others of the four in
// turn, each against an open-addressing table in LDS (medium and huge ones
// are skipped).
constexpr uint32_t kTinyBucket = 16;

__global__ __launch_bounds__(kBlock) void k_dedup_small(uint64_t n_next, const uint64_t *seg_off,
                                                       const uint32_t *bucket, uint32_t *out, uint32_t *uniq) {
  __shared__ uint32_t tabs[kBlock / kWave][kWaveTab];
  const int w = threadIdx.x / kWave;
  const uint32_t ln = threadIdx.x & 63, grp = ln / kTinyBucket, gl = ln % kTinyBucket;
  uint32_t *tab = tabs[w];
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / kWave);
  for (uint64_t q = static_cast<uint64_t>(blockIdx.x) * (kBlock / kWave) + w; 4 * q < n_next; q += waves) {
    const uint64_t ag = 4 * q + grp;  // this lane group's bucket
    uint64_t gbase = 0;
    uint32_t gs = 0;
    if (ag < n_next) {
      gbase = seg_off[ag];
      gs = static_cast<uint32_t>(seg_off[ag + 1] - gbase);
    }
    const bool tiny = gs <= kTinyBucket;
    const uint64_t big = __ballot(!tiny);
    if (big != ~0ull) {  // the tiny buckets of the four, by their lane groups
      const uint32_t v = tiny && gl < gs ? bucket[gbase + gl] : kNone;
      bool keep = tiny && gl < gs;
      for (uint32_t j = 0; j < kTinyBucket; ++j) {
        const uint32_t u = static_cast<uint32_t>(__shfl(static_cast<int>(v), static_cast<int>(grp * kTinyBucket + j)));
        if (j < gl && u == v) keep = false;
      }
      const uint32_t gm = static_cast<uint32_t>((__ballot(keep) >> (grp * kTinyBucket)) & 0xFFFFu);
      if (keep) out[gbase + __popc(gm & ((1u << gl) - 1u))] = v;
      if (tiny && gl == 0 && ag < n_next) uniq[ag] = __popc(gm);
    }
    for (uint32_t k = 0; k < 4; ++k) {  // the others in turn, by the whole wave
      if (!((big >> (k * kTinyBucket)) & 1ull)) continue;
      const uint64_t a2 = 4 * q + k;
      const uint64_t base = __shfl(gbase, static_cast<int>(k * kTinyBucket));
      const uint32_t s = static_cast<uint32_t>(__shfl(static_cast<int>(gs), static_cast<int>(k * kTinyBucket)));
      if (s > kDedupWaveCap) continue;  // listed as medium or huge
      if (s <= 1) {
        if (ln == 0) {
          uniq[a2] = s;
          if (s) out[base] = bucket[base];
        }
        continue;
      }
      uint32_t tsz = 2;
      while (tsz < 2u * s) tsz <<= 1;
      const uint32_t mask = tsz - 1;
      for (uint32_t j = ln; j < tsz; j += kWave) tab[j] = kNone;
      wave_fence();
      for (uint32_t j = ln; j < s; j += kWave) lds_insert(tab, mask, bucket[base + j]);
      wave_fence();
      // emit table contents in slot order
      uint32_t written = 0;
      for (uint32_t j0 = 0; j0 < tsz; j0 += kWave) {
        const uint32_t j = j0 + ln;
        const uint32_t val = j < tsz ? tab[j] : kNone;
        const bool keep = val != kNone;
        const uint64_t m = __ballot(keep);
        if (keep) out[base + written + __popcll(m & lanes_below())] = val;
        written += __popcll(m);
      }
      if (ln == 0) uniq[a2] = written;
      wave_fence();
    }
  }
}

// Exclusive block-wide scan of v over B threads (wave scans, one LDS pass
// over the wave sums); total = the block's sum.  Every thread calls it.
template <int B>
__device__ __forceinline__ uint32_t block_scan_waves(uint32_t v, uint32_t *wsum, uint32_t &total) {
  const uint32_t ln = lane_id(), wv = threadIdx.x / kWave;
  uint32_t inc = v;
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o);
    if (ln >= static_cast<uint32_t>(o)) inc += t;
  }
  if (ln == kWave - 1) wsum[wv] = inc;
  __syncthreads();
  uint32_t before = 0;
  total = 0;
  for (uint32_t i = 0; i < B / kWave; ++i) {
    const uint32_t x = wsum[i];
    before += i < wv ? x : 0u;
    total += x;
  }
  __syncthreads();  // wsum is free again
  return before + inc - v;
}

// Medium buckets with s_lo <= size <= s_hi, one workgroup of B threads each,
// against a table of Tab LDS slots (at least twice the largest size): the
// smaller class takes a quarter of the LDS, so four times as many workgroups
// share a CU; the larger class, two workgroups to a CU, takes 1024 threads.
template <uint32_t Tab, int B>
__global__ __launch_bounds__(B) void k_dedup_medium(uint32_t n_medium, const uint32_t *medium,
                                                   const uint64_t *seg_off, const uint32_t *bucket, uint32_t *out,
                                                   uint32_t *uniq, uint32_t s_lo, uint32_t s_hi) {
  __shared__ uint32_t tab[Tab];
  __shared__ uint32_t wsum[B / kWave];
  const int tid = threadIdx.x;
  for (uint32_t i = blockIdx.x; i < n_medium; i += gridDim.x) {
    const uint32_t a2 = medium[i];
    const uint64_t base = seg_off[a2];
    const uint32_t s = static_cast<uint32_t>(seg_off[a2 + 1] - base);
    if (s < s_lo || s > s_hi) continue;  // the other class's launch takes it
    uint32_t tsz = 2;
    while (tsz < 2u * s) tsz <<= 1;
    const uint32_t mask = tsz - 1;
    for (uint32_t j = tid; j < tsz; j += B) tab[j] = kNone;
    __syncthreads();
    for (uint32_t j = tid; j < s; j += B) lds_insert(tab, mask, bucket[base + j]);
    __syncthreads();
    // each thread owns the slots tid, tid + B, ... (consecutive lanes read
    // consecutive words: no LDS bank conflicts) and writes its kept values as
    // one run, the runs in thread order
    uint32_t c = 0;
    for (uint32_t j = tid; j < tsz; j += B) c += tab[j] != kNone;
    uint32_t tot;
    uint32_t p = block_scan_waves<B>(c, wsum, tot);
    for (uint32_t j = tid; j < tsz; j += B)
      if (tab[j] != kNone) out[base + p++] = tab[j];
    if (tid == 0) uniq[a2] = tot;
    __syncthreads();
  }
}

// Huge buckets: one global open-addressing table per bucket (pre-filled with
// kNone), processed in fixed-size chunks so that every block has the same work.
__global__ __launch_bounds__(kBlock) void k_dedup_huge_insert(const HugeChunk *chunks, uint32_t n_chunks,
                                                             uint32_t *tables, const uint32_t *bucket) {
  for (uint32_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    const HugeChunk h = chunks[c];
    uint32_t *tab = tables + h.tab;
    const uint32_t mask = h.tab_mask;
    for (uint64_t j = h.lo + threadIdx.x; j < h.hi; j += kBlock) {
      const uint32_t b = bucket[h.base + j];
      uint32_t k = hash32(b) & mask;
      while (true) {
        // a slot never changes once filled, so a non-empty read is exact
        const uint32_t cur = tab[k];
        if (cur == b) break;
        if (cur == kNone) {
          const uint32_t prev = atomicCAS(&tab[k], kNone, b);
          if (prev == kNone || prev == b) break;
        }
        k = (k + 1) & mask;
      }
    }
  }
}

// Table slots [lo, hi) of a huge bucket.  Each wave reads its slots in
// coalesced 64-slot rows twice: first to count the kept values (one atomic per
// block reserves the output range), then to write them with ballot prefixes.
__global__ __launch_bounds__(kBlock) void k_dedup_huge_emit(const HugeChunk *chunks, uint32_t n_chunks,
                                                           const uint32_t *tables, uint32_t *dst, uint32_t *hcount) {
  __shared__ uint32_t wave_tot[kBlock / kWave];
  __shared__ uint32_t at;
  const int w = threadIdx.x / kWave;
  const uint32_t ln = lane_id();
  for (uint32_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    const HugeChunk h = chunks[c];
    const uint32_t *tab = tables + h.tab;
    uint32_t cnt = 0;
    for (uint64_t r = h.lo + static_cast<uint64_t>(w) * kWave; r < h.hi; r += kBlock) {
      const uint64_t j = r + ln;
      cnt += __popcll(__ballot(j < h.hi && tab[j] != kNone));
    }
    if (ln == 0) wave_tot[w] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t tot = wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
      at = tot ? atomicAdd(&hcount[h.huge], tot) : 0u;
    }
    __syncthreads();
    uint64_t out = at;
    for (int w2 = 0; w2 < w; ++w2) out += wave_tot[w2];
    for (uint64_t r = h.lo + static_cast<uint64_t>(w) * kWave; r < h.hi; r += kBlock) {
      const uint64_t j = r + ln;
      const uint32_t val = j < h.hi ? tab[j] : kNone;
      const bool keep = val != kNone;
      const uint64_t m = __ballot(keep);
      if (keep) dst[h.base + out + __popcll(m & lanes_below())] = val;
      out += __popcll(m);
    }
    __syncthreads();
  }
}

// Huge buckets against per-bucket bitmaps over the label space: the first
// lane to set a value's bit keeps it (read before the atomic: a set bit is
// never stale, bits only get set).  The chunks are numbered bucket after
// bucket (cstart[h] = first chunk of huge bucket h, *total = all chunks) and
// found by binary search; they are spread over the whole grid, so few blocks
// work on one bitmap word at a time (the device performs these atomics beyond
// the per-XCD caches).  A block first drops the repeats inside its chunk
// through a direct-mapped LDS filter -- a value is dropped there only when a
// lane of the block holding the same value went on to the bitmap, so exactly
// one copy of every value is kept per bucket -- then stages the kept values in
// LDS and appends them to the bucket's run in out, reserving the range with
// one atomic on uniq (zeroed for the huge buckets).
constexpr int kHugeFilter = 4096;
constexpr uint32_t kStagePiece = 4096;  // entries of a chunk staged between two appends
constexpr int kMarkBlock = 1024;       // threads of a mark workgroup
constexpr int kMarkFilter = 16384;     // LDS filter entries (64 KB)
constexpr uint64_t kMapChunk = 65536;  // elements per map-bucket chunk

__global__ __launch_bounds__(kBlock) void k_dedup_huge_bits(uint32_t n_huge, const uint32_t *huge,
                                                           const uint64_t *seg_off, const uint64_t *cstart,
                                                           const uint64_t *total, uint32_t *bitmaps, uint64_t words,
                                                           const uint32_t *bucket, uint32_t *out, uint32_t *uniq) {
  __shared__ uint32_t filt[kHugeFilter];
  __shared__ uint32_t stage[kStagePiece];
  __shared__ uint32_t n_kept, at;
  const uint64_t T = *total;
  for (uint64_t c = blockIdx.x; c < T; c += gridDim.x) {
    uint32_t lo = 0, hi = n_huge;  // cstart[lo] <= c < cstart[hi] (cstart[n_huge] = T)
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (cstart[mid] <= c)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t a2 = huge[lo];
    const uint64_t base = seg_off[a2], size = seg_off[a2 + 1] - base;
    const uint64_t e_lo = (c - cstart[lo]) * kBitsChunk, e_hi = min(size, e_lo + kBitsChunk);
    uint32_t *bm = bitmaps + static_cast<uint64_t>(lo) * words;
    for (int j = threadIdx.x; j < kHugeFilter; j += kBlock) filt[j] = kNone;
    for (uint64_t p0 = e_lo; p0 < e_hi; p0 += kStagePiece) {
      const uint64_t p1 = min(e_hi, p0 + kStagePiece);
      if (threadIdx.x == 0) n_kept = 0;
      __syncthreads();
      // four elements per lane at a time: their loads, bit reads and
      // atomics each in flight together (the filter is probed in order)
      for (uint64_t jb = p0 + threadIdx.x; jb < p1; jb += 4 * kBlock) {
        uint32_t vs[4], bit[4];
        uint32_t *w[4];
        bool keep[4], probe[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) vs[u] = jb + u * kBlock < p1 ? bucket[base + jb + u * kBlock] : 0u;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t v = vs[u];
          keep[u] = probe[u] = false;
          bit[u] = 1u << (v & 31u);
          w[u] = &bm[(v >> 5) < words ? (v >> 5) : 0];
          if (jb + u * kBlock >= p1) continue;
          if ((v >> 5) >= words) {  // outside the label space: kept unchecked (never the case)
            keep[u] = true;
            continue;
          }
          const uint32_t f = hash32(v) & (kHugeFilter - 1);
          if (filt[f] != v) {
            filt[f] = v;
            probe[u] = true;
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) probe[u] = probe[u] && (*w[u] & bit[u]) == 0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (probe[u]) keep[u] = (atomicOr(w[u], bit[u]) & bit[u]) == 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint64_t m = __ballot(keep[u]);
          uint32_t wbase = 0;
          if (lane_id() == 0 && m) wbase = atomicAdd(&n_kept, static_cast<uint32_t>(__popcll(m)));
          wbase = __shfl(wbase, 0);
          if (keep[u]) stage[wbase + __popcll(m & lanes_below())] = vs[u];
        }
      }
      __syncthreads();
      const uint32_t k = n_kept;
      if (threadIdx.x == 0) at = k ? atomicAdd(&uniq[a2], k) : 0u;
      __syncthreads();
      for (uint32_t j = threadIdx.x; j < k; j += kBlock) out[base + at + j] = stage[j];
      __syncthreads();
    }
  }
}

// Chunks of huge bucket h when its size lies in [lo, hi), else none.
struct HugeChunkLoad {
  const uint32_t *huge;
  const uint64_t *seg_off;
  uint64_t lo, hi;
  __device__ uint64_t operator()(uint64_t h) const {
    const uint32_t a2 = huge[h];
    const uint64_t size = seg_off[a2 + 1] - seg_off[a2];
    return size >= lo && size < hi ? (size + kBitsChunk - 1) / kBitsChunk : 0;
  }
};

// (chunks, is a map bucket) of huge bucket h: map buckets are those of at
// least map_min elements.
struct MapChunkLoad {
  const uint32_t *huge;
  const uint64_t *seg_off;
  uint64_t map_min;
  __device__ Pair64 operator()(uint64_t h) const {
    const uint32_t a2 = huge[h];
    const uint64_t size = seg_off[a2 + 1] - seg_off[a2];
    return size >= map_min ? Pair64{(size + kMapChunk - 1) / kMapChunk, 1} : Pair64{0, 0};
  }
};

// cstart[h] = first map chunk, mslot[h] = map slot (of map buckets; the next
// slot for the others), mlist[slot] = h.
struct MapChunkStore {
  uint64_t *cstart;
  uint32_t *mslot, *mlist;
  __device__ void operator()(uint64_t h, Pair64 prefix, Pair64 v) const {
    cstart[h] = prefix.a;
    mslot[h] = static_cast<uint32_t>(prefix.b);
    if (v.b) mlist[prefix.b] = static_cast<uint32_t>(h);
  }
};

// Huge buckets of at least map_min elements against byte maps over the label
// space.  Every element stores a 1 at its value's byte: repeats store the same
// byte, so no lane needs to learn whether it came first and the device
// performs no atomics (the bitmap path above needs one per distinct value).
// The direct-mapped LDS filter only saves the stores of values repeated within
// a chunk.  Map buckets take map slots in list order (mslot[h], at most
// n_maps of them); a slot's map is `stride` bytes (the label space rounded up
// to whole emit pieces) and is all zero on entry.  The map buckets' chunks of
// kMapChunk elements are numbered as for the bitmaps (cstart, *total); large
// workgroups and chunks give the filter more repeats to catch.
__global__ __launch_bounds__(kMarkBlock) void k_dedup_huge_mark(uint32_t n_huge, const uint32_t *huge,
                                                               const uint64_t *seg_off, const uint64_t *cstart,
                                                               const uint64_t *total, const uint32_t *mslot,
                                                               uint64_t n_maps, uint8_t *maps, uint64_t stride,
                                                               uint64_t space, const uint32_t *bucket) {
  __shared__ uint32_t filt[kMarkFilter];
  const uint64_t T = *total;
  for (uint64_t c = blockIdx.x; c < T; c += gridDim.x) {
    uint32_t lo = 0, hi = n_huge;  // cstart[lo] <= c < cstart[hi] (cstart[n_huge] = T)
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (cstart[mid] <= c)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t a2 = huge[lo], slot = mslot[lo];
    if (slot >= n_maps) continue;  // the caller's bound on the map buckets holds (never the case)
    const uint64_t base = seg_off[a2], size = seg_off[a2 + 1] - base;
    const uint64_t e_lo = (c - cstart[lo]) * kMapChunk, e_hi = min(size, e_lo + kMapChunk);
    uint8_t *bm = maps + static_cast<uint64_t>(slot) * stride;
    __syncthreads();  // the previous chunk's filter probes are done
    for (int j = threadIdx.x; j < kMarkFilter; j += kMarkBlock) filt[j] = kNone;
    __syncthreads();
    for (uint64_t jb = e_lo + threadIdx.x; jb < e_hi; jb += 4 * kMarkBlock) {  // four loads in flight per lane
      uint32_t vs[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) vs[u] = jb + u * kMarkBlock < e_hi ? bucket[base + jb + u * kMarkBlock] : kNone;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t v = vs[u];
        if (v >= space) continue;  // past the chunk, or (never) outside the label space
        const uint32_t f = hash32(v) & (kMarkFilter - 1);
        if (filt[f] != v) {
          filt[f] = v;
          bm[v] = 1;
        }
      }
    }
  }
}

// The values whose byte is set, appended to their bucket's run in out: each
// block step covers one piece of one map (16 bytes per thread), counts the set
// bytes with wave scans, reserves the piece's range with one atomic on uniq
// (zeroed for the huge buckets) and writes the values in map order.  Every
// nonzero 16 bytes read are cleared again, so the maps end all zero.
// mlist[k] = list index of the bucket in map slot k; *n_used = slots in use.
constexpr uint32_t kMapPiece = 16 * kBlock;

__global__ __launch_bounds__(kBlock) void k_dedup_huge_emit_map(const uint32_t *huge, const uint64_t *seg_off,
                                                               const uint32_t *mlist, const uint64_t *n_used,
                                                               uint64_t n_maps, uint8_t *maps, uint64_t stride,
                                                               uint32_t *out, uint32_t *uniq) {
  __shared__ uint32_t wave_sum[kBlock / 64];
  __shared__ uint32_t at;
  const uint64_t per = stride / kMapPiece;
  const uint64_t T = min(*n_used, n_maps) * per;
  const uint32_t ln = lane_id(), wv = threadIdx.x / 64;
  for (uint64_t c = blockIdx.x; c < T; c += gridDim.x) {
    const uint64_t slot = c / per;
    const uint64_t p = (c - slot * per) * kMapPiece + 16ull * threadIdx.x;
    uint4 *src = reinterpret_cast<uint4 *>(maps + slot * stride + p);
    const uint4 w = *src;
    const uint32_t k = __popc(w.x) + __popc(w.y) + __popc(w.z) + __popc(w.w);  // set bytes hold 1
    if (k) *src = make_uint4(0u, 0u, 0u, 0u);
    uint32_t inc = k;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(inc, o);
      if (ln >= static_cast<uint32_t>(o)) inc += t;
    }
    if (ln == 63) wave_sum[wv] = inc;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (int i = 0; i < kBlock / 64; ++i) {
      before += i < static_cast<int>(wv) ? wave_sum[i] : 0u;
      tot += wave_sum[i];
    }
    const uint32_t a2 = huge[mlist[slot]];
    if (threadIdx.x == 0) at = tot ? atomicAdd(&uniq[a2], tot) : 0u;
    __syncthreads();
    if (k) {
      uint32_t *dst = out + seg_off[a2] + at + before + inc - k;
      const uint32_t word[4] = {w.x, w.y, w.z, w.w};
      uint32_t q = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        for (uint32_t x = word[i]; x; x &= x - 1)
          dst[q++] = static_cast<uint32_t>(p) + 4u * i + (static_cast<uint32_t>(__ffs(x) - 1) >> 3);
      }
    }
    __syncthreads();  // wave_sum and at are rewritten by the next step
  }
}

__global__ void k_zero_huge_uniq(uint32_t n_huge, const uint32_t *huge, uint32_t *uniq) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h < n_huge) uniq[huge[h]] = 0;
}

// (offset, size) of every huge bucket, for the host to plan the chunks.
__global__ void k_huge_sizes(uint32_t n_huge, const uint32_t *huge, const uint64_t *seg_off, uint64_t *out) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= n_huge) return;
  const uint32_t a2 = huge[h];
  out[2 * h] = seg_off[a2];
  out[2 * h + 1] = seg_off[a2 + 1] - seg_off[a2];
}

__global__ void k_huge_counts(uint32_t n_huge, const uint32_t *huge, const uint32_t *hcount, uint32_t *uniq) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h < n_huge) uniq[huge[h]] = hcount[h];
}

// ---------------------------------------------------------------------------
// Next-level CSR.
struct UniqLoad {
  const uint32_t *u;
  __device__ uint64_t operator()(uint64_t i) const { return u[i]; }
};
struct OffStore {
  uint64_t *off;
  uint64_t n;
  __device__ void operator()(uint64_t i, uint64_t prefix, uint64_t v) const {
    off[i] = prefix;
    if (i + 1 == n) off[n] = prefix + v;
  }
};
// Next-level columns: row a2 of the CSR copies the unique run of bucket a2,
// mapped to next-level indices.  Rows of small buckets (at most the wave cap
// of elements, so runs that short): 16 lanes to a row, four rows per wave
// step, every row's copy contiguous.  The listed (medium and huge) buckets'
// rows go in chunks of kRowChunk entries over the whole grid.
constexpr uint64_t kRowChunk = 4096;

__global__ __launch_bounds__(kBlock) void k_gather_rows(uint64_t n_next, const uint64_t *off_next,
                                                       const uint64_t *seg_off, const uint32_t *vals,
                                                       const uint32_t *map, uint32_t *col_next) {
  const uint32_t ln = lane_id(), grp = ln / 16, gl = ln % 16;
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / kWave);
  for (uint64_t q = static_cast<uint64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; 4 * q < n_next;
       q += waves) {
    const uint64_t a2 = 4 * q + grp;
    if (a2 >= n_next) continue;
    const uint64_t s0 = seg_off[a2];
    if (seg_off[a2 + 1] - s0 > kDedupWaveCap) continue;  // listed: copied in chunks
    const uint64_t o0 = off_next[a2], len = off_next[a2 + 1] - o0;
    // four values per lane in flight, then their next-level ids
    uint64_t j = gl;
    for (; j + 48 < len; j += 64) {
      uint32_t v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) v[k] = vals[s0 + j + 16 * k];
      if (map) {
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = map[v[k]];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) col_next[o0 + j + 16 * k] = v[k];
    }
    for (; j < len; j += 16) {
      const uint32_t v = vals[s0 + j];
      col_next[o0 + j] = map ? map[v] : v;
    }
  }
}

struct ListedRow {
  const uint32_t *medium, *huge;
  uint64_t n_medium;
  __device__ uint32_t operator()(uint64_t h) const { return h < n_medium ? medium[h] : huge[h - n_medium]; }
};
struct ListedChunks {
  ListedRow row;
  const uint64_t *off_next;
  __device__ uint64_t operator()(uint64_t h) const {
    const uint32_t a2 = row(h);
    return (off_next[a2 + 1] - off_next[a2] + kRowChunk - 1) / kRowChunk;
  }
};

__global__ __launch_bounds__(kBlock) void k_gather_listed(uint32_t n_list, ListedRow row, const uint64_t *cstart,
                                                         const uint64_t *total, const uint64_t *off_next,
                                                         const uint64_t *seg_off, const uint32_t *vals,
                                                         const uint32_t *map, uint32_t *col_next) {
  const uint64_t T = *total;
  for (uint64_t c = blockIdx.x; c < T; c += gridDim.x) {
    uint32_t lo = 0, hi = n_list;  // cstart[lo] <= c < cstart[hi] (cstart[n_list] = T)
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (cstart[mid] <= c)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t a2 = row(lo);
    const uint64_t o0 = off_next[a2], len = off_next[a2 + 1] - o0, s0 = seg_off[a2];
    const uint64_t j0 = (c - cstart[lo]) * kRowChunk, j1 = min(len, j0 + kRowChunk);
    for (uint64_t jb = j0 + threadIdx.x; jb < j1; jb += 4 * kBlock) {  // four loads in flight per lane
      uint32_t v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = jb + u * kBlock < j1 ? vals[s0 + jb + u * kBlock] : 0u;
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (jb + u * kBlock < j1) col_next[o0 + jb + u * kBlock] = map ? map[v[u]] : v[u];
    }
  }
}

// ---------------------------------------------------------------------------
// Undo and canonical labels.
// Grid-stride, four vertices per lane per step, so that their chains of
// dependent loads (label, next-level index, representative) overlap.
__global__ __launch_bounds__(kBlock) void k_undo(uint64_t n, const uint32_t *lab, const uint32_t *newidx,
                                                const uint32_t *gid, const uint32_t *rep_up, uint32_t *rep_out) {
  constexpr int kU = 4;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
  for (uint64_t v0 = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; v0 < n; v0 += kU * stride) {
    uint32_t a[kU], a2[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) a[u] = v0 + u * stride < n ? lab[v0 + u * stride] : 0u;
#pragma unroll
    for (int u = 0; u < kU; ++u) a2[u] = v0 + u * stride < n ? newidx[a[u]] : kNone;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const uint64_t v = v0 + u * stride;
      if (v < n) rep_out[v] = a2[u] != kNone ? rep_up[a2[u]] : (gid ? gid[a[u]] : a[u]);
    }
  }
}

__global__ void k_rep_from_keys(uint64_t n, const uint64_t *key, const uint32_t *gid, uint32_t *rep) {
  uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (v >= n) return;
  const uint32_t a = static_cast<uint32_t>(key[v]);
  rep[v] = gid ? gid[a] : a;
}

// minid[rep[v]] = min over v of the input id.  Grid-stride: each wave merges
// the lanes sharing its first lane's representative and carries that minimum
// while the next steps keep the same representative (the giant component's),
// so one atomic per wave serves it instead of one per step; atomics on one
// address are serialised by the device.  Other lanes take their own atomic
// (skipped when minid is already smaller).
__global__ __launch_bounds__(kBlock) void k_canon_min(uint64_t n, const uint32_t *rep, const uint32_t *gid,
                                                     uint32_t *minid) {
  uint32_t carry_key = kNone, carry_min = kNone;  // wave-uniform
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
  const uint64_t v0 = blockIdx.x * static_cast<uint64_t>(kBlock) + (threadIdx.x & ~(kWave - 1));
  for (uint64_t w = v0; w < n; w += stride) {  // w: the wave's first vertex this step
    const uint64_t v = w + lane_id();
    const bool live = v < n;
    const uint32_t key = live ? rep[v] : 0u;
    const uint32_t id = live ? (gid ? gid[v] : static_cast<uint32_t>(v)) : 0u;
    const uint32_t lk = __shfl(key, 0);  // lane 0 is live whenever the wave has a step
    const bool same = live && key == lk;
    uint32_t mn = same ? id : kNone;
    for (int o = 32; o > 0; o >>= 1) mn = min(mn, static_cast<uint32_t>(__shfl_xor(mn, o)));
    if (lk == carry_key) {
      carry_min = min(carry_min, mn);
    } else {
      if (carry_key != kNone && lane_id() == 0 && carry_min < minid[carry_key]) atomicMin(&minid[carry_key], carry_min);
      carry_key = lk;
      carry_min = mn;
    }
    if (live && !same && id < minid[key]) atomicMin(&minid[key], id);
  }
  if (carry_key != kNone && lane_id() == 0 && carry_min < minid[carry_key]) atomicMin(&minid[carry_key], carry_min);
}

// Block sum of the threads' counts c, written to partial[blockIdx.x].
__device__ __forceinline__ void block_sum(uint32_t c, uint32_t *partial) {
  __shared__ uint32_t wc[kBlock / kWave];
  for (int o = 32; o > 0; o >>= 1) c += static_cast<uint32_t>(__shfl_xor(static_cast<int>(c), o));
  if (lane_id() == 0) wc[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

// Grid-stride (a bounded grid keeps the partial counts few for the one-block
// sum that follows).
constexpr uint32_t kCountBlocks = 2048;

__global__ __launch_bounds__(kBlock) void k_canon_label(uint64_t n, const uint32_t *rep, const uint32_t *gid,
                                                       const uint32_t *minid, uint32_t *labels, uint32_t *partial) {
  uint32_t roots = 0;
  // four vertices per lane per step: their representatives, then the
  // representatives' minima, in flight together
  constexpr int kU = 4;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
  for (uint64_t v0 = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; v0 < n; v0 += kU * stride) {
    uint32_t r[kU], id[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const uint64_t v = v0 + u * stride;
      r[u] = v < n ? rep[v] : 0u;
      id[u] = v < n ? (gid ? gid[v] : static_cast<uint32_t>(v)) : 0u;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) r[u] = v0 + u * stride < n ? minid[r[u]] : 0u;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (v0 + u * stride >= n) continue;
      labels[id[u]] = r[u];
      roots += r[u] == id[u];
    }
  }
  block_sum(roots, partial);
}

// Level 0 without isolated vertices: newid[v] (kNone for degree 0), gid of
// every kept vertex, its row offset, and the columns mapped through newid
// (rows keep their order, so the edge order is unchanged).
struct IsoLoad {
  const uint64_t *off;
  __device__ uint64_t operator()(uint64_t v) const { return off[v + 1] > off[v] ? 1ull : 0ull; }
};
struct IsoStore {
  const uint64_t *off;
  uint32_t *newid, *gid0;
  uint64_t *off_c;
  __device__ void operator()(uint64_t v, uint64_t prefix, uint64_t keep) const {
    newid[v] = keep ? static_cast<uint32_t>(prefix) : kNone;
    if (keep) {
      gid0[prefix] = static_cast<uint32_t>(v);
      off_c[prefix] = off[v];
    }
  }
};
__global__ void k_map_cols(uint64_t m2, const uint32_t *col, const uint32_t *newid, uint32_t *out) {
  for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < m2;
       e += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    out[e] = newid[col[e]];
}
__global__ void k_iso_finish(const uint64_t *total, uint64_t m2, uint64_t *off_c, uint64_t *n_out) {
  off_c[*total] = m2;
  *n_out = *total;
}
__global__ void k_identity(uint64_t n, uint32_t *p) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = static_cast<uint32_t>(i);
}

__global__ __launch_bounds__(kBlock) void k_labels_from_keys(uint64_t n, const uint64_t *key, uint32_t *labels,
                                                            uint32_t *partial) {
  uint32_t roots = 0;
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const uint32_t l = static_cast<uint32_t>(key[v]);
    labels[v] = l;
    roots += l == static_cast<uint32_t>(v);
  }
  block_sum(roots, partial);
}

// out = sum of partial[0..count) (one block).
__global__ __launch_bounds__(kBlock) void k_reduce_sum(const uint32_t *partial, uint64_t count, uint32_t *out) {
  __shared__ uint64_t smem[kBlock];
  uint64_t acc = 0;
  for (uint64_t i = threadIdx.x; i < count; i += kBlock) acc += partial[i];
  uint64_t tot;
  block_exclusive_scan_shift<uint64_t>(acc, smem, &tot);
  if (threadIdx.x == 0) *out = static_cast<uint32_t>(tot);
}


// ---------------------------------------------------------------------------
// METIS text parsed on the device (ReadMETISFile, lib/io/graph_io.h:281-382,
// ParseVertexFilestream :1207-1248): the bytes after the header line are
// split at '\n'; lines starting with '%' are comments; the k-th other line
// lists the 1-based neighbours of vertex k, for k < n.
__device__ __forceinline__ bool byte_digit(uint8_t c) { return c >= '0' && c <= '9'; }
__device__ __forceinline__ bool byte_blank(uint8_t c) {
  return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f' || c == '\n';
}

struct ByteClassLoad {
  const uint8_t *b;
  int what;  // kBytesNewline, kBytesToken (first digit of a number), kBytesOther (neither digit nor blank)
  __device__ uint64_t operator()(uint64_t i) const {
    const uint8_t c = b[i];
    if (what == kBytesNewline) return c == '\n' ? 1ull : 0ull;
    if (what == kBytesToken) return byte_digit(c) && (i == 0 || !byte_digit(b[i - 1])) ? 1ull : 0ull;
    return !byte_digit(c) && !byte_blank(c) ? 1ull : 0ull;
  }
};
struct PosStore {
  uint64_t *pos;  // null: count only
  __device__ void operator()(uint64_t i, uint64_t prefix, uint64_t v) const {
    if (v && pos) pos[prefix] = i;
  }
};

// Line j starts after the j-th newline; it is a vertex line unless its first
// byte is '%' (an empty line is a vertex without neighbours).
struct VertexLineLoad {
  const uint8_t *b;
  const uint64_t *nlpos;
  uint64_t bytes;
  __device__ uint64_t operator()(uint64_t j) const {
    const uint64_t st = j ? nlpos[j - 1] + 1 : 0;
    return st < bytes && b[st] == '%' ? 0ull : 1ull;
  }
};
struct VertexLineStore {
  uint32_t *vline;
  uint64_t n;
  __device__ void operator()(uint64_t j, uint64_t prefix, uint64_t v) const {
    vline[j] = v && prefix < n ? static_cast<uint32_t>(prefix) : kNone;
  }
};

// Line of byte position pos: the number of newlines before it.
__device__ __forceinline__ uint64_t line_of(const uint64_t *nlpos, uint64_t nl, uint64_t pos) {
  uint64_t lo = 0, hi = nl;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (nlpos[mid] < pos)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Every number: its vertex (kNone outside the first n vertex lines) and its
// value; *bad |= 1 for a value of 0, above n or of more than 10 digits on a
// vertex line.
__global__ void k_metis_tokens(const uint8_t *b, uint64_t bytes, const uint64_t *tokpos, uint64_t count,
                               const uint64_t *nlpos, uint64_t nl, const uint32_t *vline, uint64_t n, uint32_t *tv,
                               uint32_t *tval, int *bad) {
  const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  bool wrong = false;
  if (t < count) {
    const uint64_t pos = tokpos[t];
    const uint32_t v = vline[line_of(nlpos, nl, pos)];
    uint64_t x = 0;
    if (v != kNone) {
      int d = 0;
      for (uint64_t p = pos; p < bytes && byte_digit(b[p]) && d < 11; ++p, ++d) x = x * 10 + (b[p] - '0');
      wrong = d > 10 || x == 0 || x > n;
    }
    tv[t] = v;
    tval[t] = wrong ? 0u : static_cast<uint32_t>(x);
  }
  if (__ballot(wrong) != 0 && lane_id() == 0) atomicOr(bad, 1);
}

// *bad |= 2 when a byte that is neither digit nor blank lies on a vertex line.
__global__ void k_metis_other(const uint64_t *otherpos, uint64_t count, const uint64_t *nlpos, uint64_t nl,
                              const uint32_t *vline, int *bad) {
  const uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  const bool wrong = k < count && vline[line_of(nlpos, nl, otherpos[k])] != kNone;
  if (__ballot(wrong) != 0 && lane_id() == 0) atomicOr(bad, 2);
}


// Edge lines of a sorted edge file (ReadSortedEdgeFile, lib/io/graph_io.h:
// 385-669, ParseEdgeFilestream :1250-1300): a line that is neither empty nor
// a comment ('%') holds "<type> u v"; a trailing '\r' is dropped.
struct EdgeLineSpan {
  const uint8_t *b;
  const uint64_t *nlpos;
  uint64_t bytes, nl;
  __device__ void operator()(uint64_t j, uint64_t &st, uint64_t &en) const {
    st = j ? nlpos[j - 1] + 1 : 0;
    en = j < nl ? nlpos[j] : bytes;
    if (en > st && b[en - 1] == '\r') --en;
  }
};
struct EdgeLineLoad {
  EdgeLineSpan span;
  __device__ uint64_t operator()(uint64_t j) const {
    uint64_t st, en;
    span(j, st, en);
    return st < en && span.b[st] != '%' ? 1ull : 0ull;
  }
};

// Edge k comes from the k-th edge line: the type token is skipped, then two
// numbers are read.  *bad |= 1 for a line without two numbers, |= 2 for an id
// outside [1, n].
__global__ void k_edge_text(EdgeLineSpan span, const uint64_t *eline, uint64_t count, uint64_t n, uint32_t *src,
                            uint32_t *dst, int *bad) {
  const uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  int wrong = 0;
  if (k < count) {
    uint64_t st, en;
    span(eline[k], st, en);
    const uint8_t *b = span.b;
    uint64_t p = st;
    while (p < en && (b[p] == ' ' || b[p] == '\t')) ++p;
    while (p < en && b[p] != ' ' && b[p] != '\t') ++p;
    uint64_t x[2] = {0, 0};
    for (int i = 0; i < 2 && !wrong; ++i) {
      while (p < en && (b[p] == ' ' || b[p] == '\t')) ++p;
      if (p >= en || !byte_digit(b[p])) wrong = 1;
      for (; !wrong && p < en && byte_digit(b[p]); ++p) {
        const unsigned d = b[p] - '0';
        if (x[i] > (~0ull - d) / 10) wrong = 1;
        x[i] = x[i] * 10 + d;
      }
    }
    if (!wrong && (x[0] == 0 || x[1] == 0 || x[0] > n || x[1] > n)) wrong = 2;
    src[k] = wrong ? 0u : static_cast<uint32_t>(x[0] - 1);
    dst[k] = wrong ? 0u : static_cast<uint32_t>(x[1] - 1);
  }
  const uint64_t m1 = __ballot(wrong == 1), m2 = __ballot(wrong == 2);
  if (lane_id() == 0 && (m1 | m2)) atomicOr(bad, (m1 ? 1 : 0) | (m2 ? 2 : 0));
}

struct EdgeKeepLoad {
  const uint32_t *tv;
  __device__ uint64_t operator()(uint64_t t) const { return tv[t] != kNone ? 1ull : 0ull; }
};
struct EdgeStore {
  const uint32_t *tv, *tval;
  uint32_t *src, *dst;
  __device__ void operator()(uint64_t t, uint64_t prefix, uint64_t v) const {
    if (v) {
      src[prefix] = tv[t];
      dst[prefix] = tval[t] - 1;
    }
  }
};

}  // namespace

// ===========================================================================
// Launch wrappers.

void fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL((k_fill<uint32_t>), dim3(grid_for(n) > 8192 ? 8192 : grid_for(n)), dim3(kBlock), 0, s, p, n, v);
  KC_LAUNCH_CHECK();
}
void fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL((k_fill<uint64_t>), dim3(grid_for(n) > 8192 ? 8192 : grid_for(n)), dim3(kBlock), 0, s, p, n, v);
  KC_LAUNCH_CHECK();
}

void init_keys_exp(uint64_t n, const uint32_t *gid, const uint32_t *idx, DeviateArgs d, LearnedArgs la, uint64_t *key,
                   hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_init_keys_exp, dim3(grid_for(n)), dim3(kBlock), 0, s, n, gid, idx, d, la, key);
  KC_LAUNCH_CHECK();
}
void init_keys_index(uint64_t n, uint64_t *key, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_init_keys_index, dim3(grid_for(n)), dim3(kBlock), 0, s, n, key);
  KC_LAUNCH_CHECK();
}
void init_keys_uniform(uint64_t n, const uint32_t *gid, const uint32_t *idx, int32_t seed, uint32_t it, uint64_t *key,
                       hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_init_keys_uniform, dim3(grid_for(n)), dim3(kBlock), 0, s, n, gid, idx, seed, it, key);
  KC_LAUNCH_CHECK();
}
void deviates_only(uint32_t gid0, uint32_t count, DeviateArgs d, uint32_t *out, hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_deviates_only, dim3(grid_for(count)), dim3(kBlock), 0, s, gid0, count, d, out);
  KC_LAUNCH_CHECK();
}

// Edge slots of the rows at distance t (the frontier of round t) and of the
// rows above it: out[0] and out[1] (zeroed by the caller); one atomic pair
// per block.
__global__ __launch_bounds__(kBlock) void k_round_volumes(uint64_t n, const uint64_t *off, const uint64_t *key,
                                                         uint32_t t, unsigned long long *out) {
  __shared__ unsigned long long part[2][kBlock / kWave];
  uint64_t at = 0, above = 0;
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const uint32_t d = static_cast<uint32_t>(key[v] >> 32);
    const uint64_t deg = off[v + 1] - off[v];
    at += d == t ? deg : 0;
    above += d > t ? deg : 0;
  }
  for (int o = 32; o > 0; o >>= 1) {
    at += static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(at), o));
    above += static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(above), o));
  }
  if (lane_id() == 0) {
    part[0][threadIdx.x / kWave] = at;
    part[1][threadIdx.x / kWave] = above;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    unsigned long long s = 0;
    for (int w = 0; w < kBlock / kWave; ++w) s += part[threadIdx.x][w];
    if (s) atomicAdd(&out[threadIdx.x], s);
  }
}

void round_volumes(uint64_t n, const uint64_t *off, const uint64_t *key, uint32_t t, uint64_t *out, hipStream_t s) {
  fill_u64(out, 2, 0, s);
  if (!n) return;
  const unsigned g = static_cast<unsigned>(std::min<uint64_t>(ceil_div(n, kBlock), 2048));
  hipLaunchKernelGGL(k_round_volumes, dim3(g), dim3(kBlock), 0, s, n, off, key, t,
                     reinterpret_cast<unsigned long long *>(out));
  KC_LAUNCH_CHECK();
}

void bin_vertices(uint64_t n, const uint64_t *off, BinCfg cfg, uint32_t *listT, uint32_t *listS, uint32_t *listW,
                  HubChunk *hubs, uint32_t *counters, uint32_t *deg_inter, uint64_t *scratch, hipStream_t s,
                  const uint64_t *key, uint32_t tmin) {
  if (!n) {
    fill_u32(counters, 4, 0, s);
    return;
  }
  const uint32_t nb = static_cast<uint32_t>(ceil_div(n, kBlock));
  uint32_t *counts = reinterpret_cast<uint32_t *>(scratch);
  uint64_t *prefix = scratch + ceil_div(4ull * nb, 2) + 8;
  uint64_t *total = prefix + 4ull * nb + 8;
  uint64_t *rest = total + 8;
  hipLaunchKernelGGL(k_bin_count, dim3(nb), dim3(kBlock), 0, s, n, off, cfg, counts, nb, key, tmin);
  KC_LAUNCH_CHECK();
  scan_exclusive<uint64_t>(4ull * nb, ArrLoadU32{counts}, ArrayStore<uint64_t>{prefix}, total, rest, s);
  hipLaunchKernelGGL(k_bin_write, dim3(nb), dim3(kBlock), 0, s, n, off, cfg, prefix, nb, listT, listS, listW, hubs,
                     deg_inter, key, tmin);
  KC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bin_totals, dim3(1), dim3(64), 0, s, prefix, total, nb, counters);
  KC_LAUNCH_CHECK();
}

void finish_bins(Bins &b) {
  const uint32_t blocksT = static_cast<uint32_t>(ceil_div(b.nT, kBlock));
  const uint32_t blocksS = static_cast<uint32_t>(ceil_div(b.nS, kBlock / 16));
  const uint32_t blocksW = static_cast<uint32_t>(ceil_div(b.nW, kBlock / kWave));
  const uint32_t blocksH = static_cast<uint32_t>(ceil_div(b.nH, kBlock / kWave));
  b.bS = blocksT;
  b.bW = b.bS + blocksS;
  b.bH = b.bW + blocksW;
  b.blocks = b.bH + blocksH;
}

void sweep(int mode, bool count, const uint64_t *off, const uint32_t *col, const uint64_t *kin, uint64_t *kout,
           uint64_t hop, const Bins &b, uint32_t *deg_inter, uint32_t *hub_inter, int *flag, const int *prev_flag,
           hipStream_t s, uint64_t block, bool hubs_first) {
  if (b.blocks == 0) return;
  SweepArgs a{off, col, kin, kout, hop, b, deg_inter, hub_inter, flag, prev_flag, block, hubs_first ? 1 : 0};
  dim3 g(b.blocks), t(kBlock);
  if (mode == kSweepInPlace && count)
    hipLaunchKernelGGL((k_sweep<kSweepInPlace, true>), g, t, 0, s, a);
  else if (mode == kSweepInPlace)
    hipLaunchKernelGGL((k_sweep<kSweepInPlace, false>), g, t, 0, s, a);
  else if (mode == kSweepJacobi)
    hipLaunchKernelGGL((k_sweep<kSweepJacobi, false>), g, t, 0, s, a);
  else if (mode == kSweepBlock)
    hipLaunchKernelGGL((k_sweep<kSweepBlock, false>), g, t, 0, s, a);
  else
    hipLaunchKernelGGL((k_sweep<kSweepCountOnly, true>), g, t, 0, s, a);
  KC_LAUNCH_CHECK();
}

void pack_keys(uint64_t n, const uint64_t *key, uint32_t *pk, uint32_t *bad, hipStream_t s) {
  fill_u32(bad, 1, 0u, s);
  if (!n) return;
  const unsigned g = static_cast<unsigned>(std::min<uint64_t>(ceil_div(n, kBlock), 4096));
  hipLaunchKernelGGL(k_pack_keys, dim3(g), dim3(kBlock), 0, s, n, key, pk, bad);
  KC_LAUNCH_CHECK();
}

void unpack_keys(uint64_t n, const uint32_t *pk, uint64_t *key, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_unpack_keys, dim3(static_cast<unsigned>(grid_for(n))), dim3(kBlock), 0, s, n, pk, key);
  KC_LAUNCH_CHECK();
}

void sweep_packed(const uint64_t *off, const uint32_t *col, uint32_t *pk, const Bins &b, int *flag,
                  const int *prev_flag, hipStream_t s, bool hubs_first) {
  if (b.blocks == 0) return;
  PackedSweepArgs a{off, col, pk, b, flag, prev_flag, hubs_first ? 1 : 0};
  hipLaunchKernelGGL(k_sweep_packed, dim3(b.blocks), dim3(kBlock), 0, s, a);
  KC_LAUNCH_CHECK();
}

void push_round(const uint64_t *off, const uint32_t *col, uint64_t *key, uint64_t hop, const Bins &b, uint32_t t,
                int *pending, const int *prev_pending, uint64_t hub_pull_deg, hipStream_t s, bool hubs_first) {
  if (b.blocks == 0) return;
  SweepArgs a{off, col, key, key, hop, b, nullptr, nullptr, pending, prev_pending};
  if (hub_pull_deg && b.nH) {
    hipLaunchKernelGGL(k_hub_pull, dim3(static_cast<unsigned>(ceil_div(static_cast<uint64_t>(b.nH) * kWave, kBlock))),
                       dim3(kBlock), 0, s, a, t, hub_pull_deg);
    KC_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_round, dim3(b.blocks), dim3(kBlock), 0, s, a, t, hubs_first ? 1 : 0);
  KC_LAUNCH_CHECK();
}

void hub_fold(const Bins &b, const uint32_t *hub_inter, uint32_t *deg_inter, uint32_t *hub_pos, uint64_t *prefix,
              uint64_t *scratch, hipStream_t s) {
  if (!b.nH) return;
  scan_exclusive<uint64_t>(b.nH, ArrLoadU32{hub_inter}, ArrayStore<uint64_t>{prefix}, prefix + b.nH, scratch, s);
  hipLaunchKernelGGL(k_hub_fold, dim3(grid_for(b.nH)), dim3(kBlock), 0, s, b.hubs, b.nH, prefix, deg_inter, hub_pos);
  KC_LAUNCH_CHECK();
}

static uint32_t row_wave_grid(uint64_t n) {
  return static_cast<uint32_t>(std::min<uint64_t>(ceil_div(n, static_cast<uint64_t>(kBlock / kWave)), 1u << 16));
}

void block_interface(const uint64_t *off, const uint32_t *col, uint64_t n, uint64_t bsize, uint64_t *key,
                     hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_block_interface, dim3(row_wave_grid(n)), dim3(kBlock), 0, s, off, col, n, bsize, key);
  KC_LAUNCH_CHECK();
}

void jump(uint64_t n, uint64_t *key, int *flag, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_jump, dim3(grid_for(n)), dim3(kBlock), 0, s, n, key, flag);
  KC_LAUNCH_CHECK();
}

uint64_t bin_scratch_words(uint64_t n) {
  const uint64_t nb = ceil_div(n, kBlock) + 1;
  return 6 * nb + 64 + scan_words_for(4 * nb, 1) + 64;
}

void huge_sizes(uint32_t n_huge, const uint32_t *huge, const uint64_t *seg_off, uint64_t *out, hipStream_t s) {
  if (!n_huge) return;
  hipLaunchKernelGGL(k_huge_sizes, dim3(static_cast<unsigned>(ceil_div(n_huge, kBlock))), dim3(kBlock), 0, s, n_huge,
                     huge, seg_off, out);
  KC_LAUNCH_CHECK();
}

void labels_radius(uint64_t n, const uint64_t *key, const uint32_t *pos, uint32_t *lab, uint32_t *radius,
                   bool with_radius, hipStream_t s, uint32_t *used) {
  if (!n) return;
  const unsigned g = static_cast<unsigned>(std::min<uint64_t>(ceil_div(n, kBlock), 2048));
  hipLaunchKernelGGL(k_labels_radius, dim3(g), dim3(kBlock), 0, s, n, key, pos, lab, used, radius,
                     with_radius ? 1 : 0);
  KC_LAUNCH_CHECK();
}

uint64_t scan_scratch_words(uint64_t n) { return scan_words_for(n, 2) + 64; }


struct RankStore {
  uint32_t *rank;
  __device__ void operator()(uint64_t i, uint64_t prefix, uint64_t v) const {
    if (v) rank[i] = static_cast<uint32_t>(prefix);
  }
};

void cluster_ranks(uint64_t n, const uint32_t *used, uint32_t *rank, uint64_t *count, uint64_t *scratch,
                   hipStream_t s) {
  scan_exclusive<uint64_t>(n, ArrLoadU32{used}, RankStore{rank}, count, scratch, s);
}

// Cluster numbering with the clusters of the first k positions first: used[a]
// = 3 for their labels (after labels_radius set 1 for every used label); a
// two-counter scan ranks those labels 0 .. H-1 and the others from H on.
__global__ void k_mark_hot(uint64_t k, const uint32_t *lab, uint32_t *used) {
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (v < k) used[lab[v]] = 3;
}
struct HotRankLoad {
  const uint32_t *used;
  __device__ Pair64 operator()(uint64_t i) const {
    const uint32_t u = used[i];
    return {u == 3 ? 1ull : 0ull, u == 1 ? 1ull : 0ull};
  }
};
struct HotRankStore {
  const uint32_t *used;
  uint32_t *rank;
  __device__ void operator()(uint64_t i, Pair64 prefix, Pair64 v) const {
    if (v.a) rank[i] = static_cast<uint32_t>(prefix.a);
    else if (v.b) rank[i] = static_cast<uint32_t>(prefix.b);  // H added by k_rank_fix
  }
};
__global__ void k_rank_fix(uint64_t n, const uint32_t *used, uint32_t *rank, const Pair64 *total, uint64_t *count) {
  const uint64_t a = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (a < n && used[a] == 1) rank[a] += static_cast<uint32_t>(total->a);
  if (a == 0) *count = total->a + total->b;
}

// Grid-stride, four vertices per lane per step (their label and rank loads
// in flight together).
__global__ __launch_bounds__(kBlock) void k_cluster_ids(uint64_t n, const uint32_t *lab, const uint32_t *rank,
                                                       uint32_t *cid) {
  constexpr int kU = 4;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
  for (uint64_t v0 = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; v0 < n; v0 += kU * stride) {
    uint32_t l[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) l[u] = v0 + u * stride < n ? lab[v0 + u * stride] : 0u;
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (v0 + u * stride < n) l[u] = rank[l[u]];
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (v0 + u * stride < n) cid[v0 + u * stride] = l[u];
  }
}

void cluster_ranks_hot(uint64_t n, uint64_t k, const uint32_t *lab, uint32_t *used, uint32_t *rank, uint64_t *count,
                       uint64_t *scratch, hipStream_t s) {
  if (!n) return;
  if (k > n) k = n;
  if (k) {
    hipLaunchKernelGGL(k_mark_hot, dim3(grid_for(k)), dim3(kBlock), 0, s, k, lab, used);
    KC_LAUNCH_CHECK();
  }
  Pair64 *total = reinterpret_cast<Pair64 *>(scratch);
  scan_exclusive<Pair64>(n, HotRankLoad{used}, HotRankStore{used, rank}, total, scratch + 4, s);
  hipLaunchKernelGGL(k_rank_fix, dim3(grid_for(n)), dim3(kBlock), 0, s, n, used, rank, total, count);
  KC_LAUNCH_CHECK();
}

void cluster_ids(uint64_t n, const uint32_t *lab, const uint32_t *rank, uint32_t *cid, hipStream_t s) {
  if (!n) return;
  const unsigned g = static_cast<unsigned>(std::min<uint64_t>(ceil_div(n, 4 * kBlock), 1u << 16));
  hipLaunchKernelGGL(k_cluster_ids, dim3(g), dim3(kBlock), 0, s, n, lab, rank, cid);
  KC_LAUNCH_CHECK();
}

void select_inter(uint64_t n, const uint64_t *key, const uint32_t *deg_inter, uint32_t *sk, uint32_t *sv,
                  uint64_t *count, uint64_t *scratch, hipStream_t s) {
  scan_exclusive<uint64_t>(n, SelectLoad{deg_inter}, SelectStore{key, sk, sv}, count, scratch, s);
}

void segments(uint64_t N, const uint32_t *sk, const uint32_t *sv, const uint32_t *deg_inter, const uint32_t *gid,
              const uint32_t *rank, uint32_t *pos_of, uint32_t *newidx, uint32_t *seg_of_cid, uint32_t *gid_next,
              uint64_t *seg_off, uint64_t *sizes, uint32_t *cache, uint64_t *scratch, hipStream_t s) {
  Pair64 *total = reinterpret_cast<Pair64 *>(scratch);
  const SegStore store{sk, sv, gid, rank, pos_of, newidx, seg_of_cid, gid_next, seg_off};
  if (ceil_div(N, kScanTile) <= 1) {  // one pass: nothing to cache
    scan_exclusive<Pair64>(N, SegLoad{sk, sv, deg_inter, cache}, store, total, scratch + 4, s);
  } else {
    const uint64_t blocks = ceil_div(N, kScanTile);
    Pair64 *sums = reinterpret_cast<Pair64 *>(scratch + 4);
    Pair64 *prefix = sums + blocks;
    uint64_t *rest = reinterpret_cast<uint64_t *>(prefix + blocks) + 4;
    hipLaunchKernelGGL((k_scan_reduce<Pair64, SegLoad>), dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, N,
                       SegLoad{sk, sv, deg_inter, cache}, sums);
    KC_LAUNCH_CHECK();
    if (blocks <= kScanFusedBlocks) {
      hipLaunchKernelGGL((k_scan_down<Pair64, SegLoadCached, SegStore>), dim3(static_cast<unsigned>(blocks)),
                         dim3(kBlock), 0, s, N, SegLoadCached{sk, cache}, store, static_cast<const Pair64 *>(nullptr),
                         total, static_cast<const Pair64 *>(sums));
    } else {
      scan_exclusive<Pair64>(blocks, ArrayLoad<Pair64>{sums}, ArrayStore<Pair64>{prefix}, total, rest, s);
      hipLaunchKernelGGL((k_scan_down<Pair64, SegLoadCached, SegStore>), dim3(static_cast<unsigned>(blocks)),
                         dim3(kBlock), 0, s, N, SegLoadCached{sk, cache}, store, static_cast<const Pair64 *>(prefix),
                         static_cast<Pair64 *>(nullptr));
    }
    KC_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_seg_finish, dim3(1), dim3(1), 0, s, total, seg_off, sizes);
  KC_LAUNCH_CHECK();
}

void scan_exclusive_u32_to_u64(uint64_t n, const uint32_t *in, uint64_t *out, uint64_t *total, uint64_t *scratch,
                               hipStream_t s) {
  scan_exclusive<uint64_t>(n, ArrLoadU32{in}, ArrayStore<uint64_t>{out}, total, scratch, s);
}

void relabel(const uint64_t *off, const uint32_t *col, const uint32_t *lab, const Bins &b, uint32_t *relab,
             uint32_t *deg_inter, uint32_t *hub_inter, uint64_t n_hot, hipStream_t s, bool hot16) {
  if (b.blocks == 0) return;
  SweepArgs a{off, col, nullptr, nullptr, 0, b, deg_inter, hub_inter, nullptr, nullptr};
  int dev = 0, cus = 0;
  KC_HIP(hipGetDevice(&dev));
  KC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const uint32_t blocks = static_cast<uint32_t>(ceil_div(b.blocks, kRelabelQuarters));
  const uint32_t grid = std::min<uint32_t>(blocks, static_cast<uint32_t>(cus > 0 ? cus : 1));
  if (hot16)
    hipLaunchKernelGGL(k_relabel_hot<true>, dim3(grid), dim3(kBlock * kRelabelQuarters), 0, s, a, lab, relab,
                       static_cast<uint32_t>(std::min<uint64_t>(n_hot, kRelabelHot16)));
  else
    hipLaunchKernelGGL(k_relabel_hot<false>, dim3(grid), dim3(kBlock * kRelabelQuarters), 0, s, a, lab, relab,
                       static_cast<uint32_t>(std::min<uint64_t>(n_hot, kRelabelHot)));
  KC_LAUNCH_CHECK();
}

void relabel_split(const uint64_t *off, const uint32_t *col, const uint32_t *lab, uint64_t n, const Bins &b,
                   const HubChunk *pieces, const uint32_t *part_list, const uint32_t *part_at, uint32_t *relab,
                   uint32_t *deg_inter, uint32_t *piece_inter, uint64_t n_hot, hipStream_t s) {
  if (b.bW == 0 && part_at[kParts + 1] == 0) return;
  SweepArgs a{off, col, nullptr, nullptr, 0, b, deg_inter, piece_inter, nullptr, nullptr};
  PartRange pr;
  for (uint32_t p = 0; p <= kParts + 1; ++p) pr.at[p] = part_at[p];
  int dev = 0, cus = 0;
  KC_HIP(hipGetDevice(&dev));
  KC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // one block per CU, a multiple of kParts (each part needs at least one)
  const uint32_t grid = std::max<uint32_t>(kParts, static_cast<uint32_t>(cus > 0 ? cus : 1) / kParts * kParts);
  hipLaunchKernelGGL(k_relabel_split, dim3(grid), dim3(kBlock * kRelabelQuarters), 0, s, a, lab, relab,
                     static_cast<uint32_t>(std::min<uint64_t>(n_hot, kRelabelHot16)), n, pieces, part_list, pr);
  KC_LAUNCH_CHECK();
}

// Bin lists reordered by level index (the key's low half: idx of the storage
// position), for round 0 of a degree-ordered level.
__global__ void k_list_keys(uint32_t c, const uint32_t *list, const uint32_t *idx, uint32_t *keys, uint32_t *vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c) return;
  const uint32_t p = list[i];
  keys[i] = idx[p];
  vals[i] = p;
}
__global__ void k_hub_keys(uint32_t c, const HubChunk *hubs, const uint32_t *idx, uint32_t *keys, uint32_t *vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c) return;
  keys[i] = idx[hubs[i].v];
  vals[i] = i;
}
__global__ void k_hub_permute(uint32_t c, const HubChunk *hubs, const uint32_t *order, HubChunk *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < c) out[i] = hubs[order[i]];
}

uint32_t order_bins_by_index(const Bins &b, const uint32_t *idx, uint64_t n, Bins &out, uint32_t *k0, uint32_t *v0,
                             uint32_t *k1, uint32_t *v1, uint64_t *scratch, hipStream_t s) {
  uint32_t median = kNone;
  int bits = 1;
  while (bits < 32 && (1ull << bits) < n) ++bits;
  const uint32_t *in[3] = {b.listT, b.listS, b.listW};
  uint32_t *dst[3] = {out.listT, out.listS, out.listW};
  const uint32_t cnt[3] = {b.nT, b.nS, b.nW};
  for (int l = 0; l < 3; ++l) {
    if (!cnt[l]) continue;
    hipLaunchKernelGGL(k_list_keys, dim3(static_cast<unsigned>(grid_for(cnt[l]))), dim3(kBlock), 0, s, cnt[l], in[l],
                       idx, k0, v0);
    KC_LAUNCH_CHECK();
    const int w = radix_sort_pairs(cnt[l], bits, k0, v0, k1, v1, scratch, s);
    KC_HIP(hipMemcpyAsync(dst[l], w ? v1 : v0, 4ull * cnt[l], hipMemcpyDeviceToDevice, s));
  }
  if (b.nH) {
    // chunks of one hub share their key, so the stable sort keeps them in order
    hipLaunchKernelGGL(k_hub_keys, dim3(static_cast<unsigned>(grid_for(b.nH))), dim3(kBlock), 0, s, b.nH, b.hubs, idx,
                       k0, v0);
    KC_LAUNCH_CHECK();
    const int w = radix_sort_pairs(b.nH, bits, k0, v0, k1, v1, scratch, s);
    hipLaunchKernelGGL(k_hub_permute, dim3(static_cast<unsigned>(grid_for(b.nH))), dim3(kBlock), 0, s, b.nH, b.hubs,
                       w ? v1 : v0, out.hubs);
    KC_LAUNCH_CHECK();
    KC_HIP(hipMemcpyAsync(&median, (w ? k1 : k0) + b.nH / 2, sizeof(median), hipMemcpyDeviceToHost, s));
    KC_HIP(hipStreamSynchronize(s));
  }
  return median;
}

void count_pieces(uint64_t n, const uint64_t *off, const uint32_t *col, uint32_t smax, uint32_t wmax, uint64_t split,
                  uint32_t *cnt, uint64_t *start, uint64_t *scratch, hipStream_t s) {
  if (!n) return;
  const unsigned g = static_cast<unsigned>(std::min<uint64_t>(grid_for(n), 65536));
  hipLaunchKernelGGL(k_piece_count, dim3(g), dim3(kBlock), 0, s, n, off, col, smax, wmax, split, cnt);
  KC_LAUNCH_CHECK();
  scan_exclusive<uint64_t>(n, ArrLoadU32{cnt}, ArrayStore<uint64_t>{start}, start + n, scratch, s);
}

void write_pieces(uint64_t n, const uint64_t *off, const uint32_t *col, uint32_t smax, uint32_t wmax, uint64_t split,
                  const uint64_t *start, HubChunk *pieces, uint8_t *part, hipStream_t s) {
  if (!n) return;
  const unsigned g = static_cast<unsigned>(std::min<uint64_t>(grid_for(n), 65536));
  hipLaunchKernelGGL(k_piece_write, dim3(g), dim3(kBlock), 0, s, n, off, col, smax, wmax, split, start, pieces,
                     part);
  KC_LAUNCH_CHECK();
}

void scatter(const uint64_t *off, const uint32_t *relab, const Bins &b, const uint32_t *deg_inter,
             const uint32_t *hub_inter, const uint32_t *hub_pos, const uint32_t *pos_of, uint32_t *bucket,
             hipStream_t s) {
  if (b.blocks == 0) return;
  ScatterArgs a{off, relab, b, deg_inter, hub_inter, hub_pos, pos_of, bucket};
  hipLaunchKernelGGL(k_scatter, dim3(b.blocks), dim3(kBlock), 0, s, a);
  KC_LAUNCH_CHECK();
}

void classify_buckets(uint64_t n, const uint64_t *n_dev, const uint64_t *seg_off, uint32_t *medium, uint32_t *huge,
                      uint64_t *counts, uint64_t *scratch, hipStream_t s) {
  scan_exclusive<Pair64>(n, ClassLoad{n_dev, n, seg_off}, ClassStore{medium, huge}, reinterpret_cast<Pair64 *>(counts),
                         scratch, s);
}

void dedup_small(uint64_t n_next, const uint64_t *seg_off, const uint32_t *bucket, uint32_t *out, uint32_t *uniq,
                 hipStream_t s) {
  if (!n_next) return;
  uint64_t blocks = ceil_div(ceil_div(n_next, 4), kBlock / kWave);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_dedup_small, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, n_next, seg_off, bucket,
                     out, uniq);
  KC_LAUNCH_CHECK();
}

void dedup_medium(uint32_t n_medium, const uint32_t *medium, const uint64_t *seg_off, const uint32_t *bucket,
                  uint32_t *out, uint32_t *uniq, hipStream_t s) {
  if (!n_medium) return;
  constexpr uint32_t kSplit = kDedupBlockCap / 4;  // sizes up to here take the small tables
  const unsigned small_blocks = n_medium > 8192 ? 8192 : n_medium, big_blocks = n_medium > 2048 ? 2048 : n_medium;
  hipLaunchKernelGGL((k_dedup_medium<2 * kSplit, kBlock>), dim3(small_blocks), dim3(kBlock), 0, s, n_medium, medium,
                     seg_off, bucket, out, uniq, 0u, kSplit);
  KC_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_dedup_medium<kBlockTab, 1024>), dim3(big_blocks), dim3(1024), 0, s, n_medium, medium, seg_off,
                     bucket, out, uniq, kSplit + 1, kDedupBlockCap);
  KC_LAUNCH_CHECK();
}

void dedup_huge(const HugeChunk *elem_chunks, uint32_t n_elem, const HugeChunk *slot_chunks, uint32_t n_slot,
                uint32_t n_huge, const uint32_t *huge, uint32_t *tables, const uint32_t *bucket, uint32_t *out,
                uint32_t *uniq, uint32_t *hcount, hipStream_t s) {
  if (!n_elem) return;
  const unsigned ge = n_elem < 8192 ? n_elem : 8192, gs = n_slot < 8192 ? n_slot : 8192;
  hipLaunchKernelGGL(k_dedup_huge_insert, dim3(ge), dim3(kBlock), 0, s, elem_chunks, n_elem, tables, bucket);
  KC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dedup_huge_emit, dim3(gs), dim3(kBlock), 0, s, slot_chunks, n_slot, tables, out, hcount);
  KC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_huge_counts, dim3(static_cast<unsigned>(ceil_div(n_huge, kBlock))), dim3(kBlock), 0, s, n_huge,
                     huge, hcount, uniq);
  KC_LAUNCH_CHECK();
}

uint64_t huge_map_stride(uint64_t space) { return (space + kMapPiece - 1) / kMapPiece * kMapPiece; }

void dedup_huge_mixed(uint32_t n_huge, const uint32_t *huge, const uint64_t *seg_off, uint64_t max_chunks,
                      uint64_t map_min, uint64_t n_maps, bool with_bits, bool zero_uniq, uint64_t *info,
                      uint32_t *bitmaps, uint64_t words, uint8_t *maps, uint64_t space, const uint32_t *bucket,
                      uint32_t *out, uint32_t *uniq, uint64_t *scratch, hipStream_t s) {
  if (!n_huge) return;
  if (!n_maps || !space) map_min = ~0ull;
  uint64_t *cbits = info, *cmap = info + n_huge + 1;
  Pair64 *mtot = reinterpret_cast<Pair64 *>(info + 2 * n_huge + 2);
  uint32_t *mslot = reinterpret_cast<uint32_t *>(info + 2 * n_huge + 4);
  uint32_t *mlist = mslot + n_huge;
  if (zero_uniq) {
    hipLaunchKernelGGL(k_zero_huge_uniq, dim3(static_cast<unsigned>(ceil_div(n_huge, kBlock))), dim3(kBlock), 0, s,
                       n_huge, huge, uniq);
    KC_LAUNCH_CHECK();
  }
  const uint64_t g = std::min<uint64_t>(8192, max_chunks);
  if (with_bits && map_min > kDedupBlockCap + 1) {  // some huge buckets against bitmaps
    scan_exclusive<uint64_t>(n_huge, HugeChunkLoad{huge, seg_off, 0, map_min}, ArrayStore<uint64_t>{cbits},
                             cbits + n_huge, scratch, s);
    fill_u32(bitmaps, static_cast<uint64_t>(n_huge) * words, 0u, s);
    hipLaunchKernelGGL(k_dedup_huge_bits, dim3(static_cast<unsigned>(g)), dim3(kBlock), 0, s, n_huge, huge, seg_off,
                       cbits, cbits + n_huge, bitmaps, words, bucket, out, uniq);
    KC_LAUNCH_CHECK();
  }
  if (map_min != ~0ull) {  // the others against byte maps
    const uint64_t stride = huge_map_stride(space);
    scan_exclusive<Pair64>(n_huge, MapChunkLoad{huge, seg_off, map_min}, MapChunkStore{cmap, mslot, mlist}, mtot,
                           scratch, s);
    hipLaunchKernelGGL(k_dedup_huge_mark, dim3(static_cast<unsigned>(g)), dim3(kMarkBlock), 0, s, n_huge, huge,
                       seg_off, cmap, &mtot->a, mslot, n_maps, maps, stride, space, bucket);
    KC_LAUNCH_CHECK();
    const uint64_t ge = std::min<uint64_t>(8192, n_maps * (stride / kMapPiece));
    hipLaunchKernelGGL(k_dedup_huge_emit_map, dim3(static_cast<unsigned>(ge)), dim3(kBlock), 0, s, huge, seg_off,
                       mlist, &mtot->b, n_maps, maps, stride, out, uniq);
    KC_LAUNCH_CHECK();
  }
}

void csr_offsets(uint64_t n_next, const uint32_t *uniq, uint64_t *off_next, uint64_t *sizes, uint64_t *scratch,
                 hipStream_t s) {
  if (n_next == 0) {
    fill_u64(off_next, 1, 0, s);
    fill_u64(sizes + 2, 1, 0, s);
    return;
  }
  uint64_t *total = scratch;
  scan_exclusive<uint64_t>(n_next, UniqLoad{uniq}, OffStore{off_next, n_next}, total, scratch + 4, s);
  hipLaunchKernelGGL(k_copy_total, dim3(1), dim3(1), 0, s, total, sizes + 2);
  KC_LAUNCH_CHECK();
}

void gather_uniques(uint64_t n_next, const uint64_t *off_next, const uint64_t *seg_off, const uint32_t *vals,
                    const uint32_t *map, const uint32_t *medium, uint64_t n_medium, const uint32_t *huge,
                    uint64_t n_huge, uint64_t *cstart, uint64_t *scratch, uint32_t *col_next, hipStream_t s) {
  if (!n_next) return;
  const uint64_t waves = ceil_div(n_next, 4);
  const uint64_t blocks = std::min<uint64_t>(ceil_div(waves, kBlock / kWave), 8192);
  hipLaunchKernelGGL(k_gather_rows, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, n_next, off_next, seg_off,
                     vals, map, col_next);
  KC_LAUNCH_CHECK();
  const uint64_t n_list = n_medium + n_huge;
  if (!n_list) return;
  const ListedRow row{medium, huge, n_medium};
  scan_exclusive<uint64_t>(n_list, ListedChunks{row, off_next}, ArrayStore<uint64_t>{cstart}, cstart + n_list, scratch,
                           s);
  hipLaunchKernelGGL(k_gather_listed, dim3(8192), dim3(kBlock), 0, s, static_cast<uint32_t>(n_list), row, cstart,
                     cstart + n_list, off_next, seg_off, vals, map, col_next);
  KC_LAUNCH_CHECK();
}

void undo(uint64_t n, const uint32_t *lab, const uint32_t *newidx, const uint32_t *gid, const uint32_t *rep_up,
          uint32_t *rep_out, hipStream_t s) {
  if (!n) return;
  const unsigned g = static_cast<unsigned>(std::min<uint64_t>(ceil_div(n, 4 * kBlock), 1u << 16));
  hipLaunchKernelGGL(k_undo, dim3(g), dim3(kBlock), 0, s, n, lab, newidx, gid, rep_up, rep_out);
  KC_LAUNCH_CHECK();
}

void rep_from_keys(uint64_t n, const uint64_t *key, const uint32_t *gid, uint32_t *rep, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_rep_from_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, n, key, gid, rep);
  KC_LAUNCH_CHECK();
}

void canon_min(uint64_t n, const uint32_t *rep, const uint32_t *gid, uint32_t *minid, hipStream_t s) {
  if (!n) return;
  const unsigned g = static_cast<unsigned>(std::min<uint64_t>(ceil_div(n, kBlock), 2048));
  hipLaunchKernelGGL(k_canon_min, dim3(g), dim3(kBlock), 0, s, n, rep, gid, minid);
  KC_LAUNCH_CHECK();
}

void canon_label(uint64_t n, const uint32_t *rep, const uint32_t *gid, const uint32_t *minid, uint32_t *labels,
                 uint32_t *roots, uint32_t *partial, hipStream_t s) {
  if (!n) return;
  const unsigned g = static_cast<unsigned>(std::min<uint64_t>(grid_for(n), kCountBlocks));
  hipLaunchKernelGGL(k_canon_label, dim3(g), dim3(kBlock), 0, s, n, rep, gid, minid, labels, partial);
  KC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kBlock), 0, s, partial, static_cast<uint64_t>(g), roots);
  KC_LAUNCH_CHECK();
}

void drop_isolated(uint64_t n, const uint64_t *off, const uint32_t *col, uint64_t m2, uint32_t *newid, uint32_t *gid0,
                   uint64_t *off_c, uint32_t *col_c, uint64_t *n_out, uint64_t *scratch, hipStream_t s) {
  if (!n) return;
  uint64_t *total = scratch;
  scan_exclusive<uint64_t>(n, IsoLoad{off}, IsoStore{off, newid, gid0, off_c}, total, scratch + 4, s);
  hipLaunchKernelGGL(k_iso_finish, dim3(1), dim3(1), 0, s, total, m2, off_c, n_out);
  KC_LAUNCH_CHECK();
  if (m2) {
    hipLaunchKernelGGL(k_map_cols, dim3(grid_for(m2) > 65536 ? 65536 : grid_for(m2)), dim3(kBlock), 0, s, m2, col, newid,
                       col_c);
    KC_LAUNCH_CHECK();
  }
}

// Component-size histogram: count[label]++ over the vertices, then
// hist[count[c]]++ over the labels with a count, then the non-zero entries of
// hist as (size, number) pairs in increasing size.
__global__ void k_count_labels(uint64_t n, const uint32_t *labels, uint32_t *count) {
  for (uint64_t base = blockIdx.x * static_cast<uint64_t>(blockDim.x); base < n;
       base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t v = base + threadIdx.x;
    const bool live = v < n;
    peeled_fetch_add(count, live ? labels[v] : 0u, 1u, live);
  }
}
__global__ void k_size_hist(uint64_t n, const uint32_t *count, uint32_t *hist) {
  for (uint64_t base = blockIdx.x * static_cast<uint64_t>(blockDim.x); base < n;
       base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t v = base + threadIdx.x;
    const uint32_t c = v < n ? count[v] : 0u;
    peeled_fetch_add(hist, c, 1u, c != 0);
  }
}
struct HistLoad {
  const uint32_t *hist;
  __device__ uint64_t operator()(uint64_t i) const { return hist[i] ? 1ull : 0ull; }
};
struct HistStore {
  const uint32_t *hist;
  uint64_t *pairs;
  __device__ void operator()(uint64_t i, uint64_t prefix, uint64_t keep) const {
    if (keep) {
      pairs[2 * prefix] = i;
      pairs[2 * prefix + 1] = hist[i];
    }
  }
};

void component_histogram(uint64_t n, const uint32_t *labels, uint32_t *count, uint32_t *hist, uint64_t *pairs,
                         uint64_t *num_pairs, uint64_t *scratch, hipStream_t s) {
  fill_u32(count, n, 0u, s);
  fill_u32(hist, n + 1, 0u, s);
  if (!n) {
    fill_u64(num_pairs, 1, 0ull, s);
    return;
  }
  const unsigned grid = grid_for(n) > 65536 ? 65536 : grid_for(n);
  hipLaunchKernelGGL(k_count_labels, dim3(grid), dim3(kBlock), 0, s, n, labels, count);
  KC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_size_hist, dim3(grid), dim3(kBlock), 0, s, n, count, hist);
  KC_LAUNCH_CHECK();
  scan_exclusive<uint64_t>(n + 1, HistLoad{hist}, HistStore{hist, pairs}, num_pairs, scratch, s);
}

// Degree-ordered storage of level 0: sort keys ~deg (stable, so ties keep
// index order), then the maps and the CSR in storage order.
__global__ void k_perm_sortkeys(uint64_t n, const uint64_t *off, uint32_t *key, uint32_t *val) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    key[i] = ~static_cast<uint32_t>(off[i + 1] - off[i]);
    val[i] = static_cast<uint32_t>(i);
  }
}
__global__ void k_perm_maps(uint64_t n, const uint32_t *idx, const uint32_t *gid, const uint64_t *off, uint32_t *pos,
                            uint32_t *gidp, uint32_t *degp) {
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < n;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t i = idx[p];
    pos[i] = static_cast<uint32_t>(p);
    gidp[p] = gid ? gid[i] : i;
    degp[p] = static_cast<uint32_t>(off[i + 1] - off[i]);
  }
}
// One 64-lane block per row: row idx[p] of (off, col) becomes row p, its
// neighbours mapped to storage positions.
__global__ __launch_bounds__(64) void k_perm_rows(uint64_t n, const uint32_t *idx, const uint64_t *off,
                                                  const uint32_t *col, const uint32_t *pos, const uint64_t *offp,
                                                  uint32_t *colp) {
  for (uint64_t p = blockIdx.x; p < n; p += gridDim.x) {
    const uint32_t i = idx[p];
    const uint64_t s0 = off[i], len = off[i + 1] - s0, d0 = offp[p];
    for (uint64_t k = threadIdx.x; k < len; k += 64) colp[d0 + k] = pos[col[s0 + k]];
  }
}

void level0_order(uint64_t n, const uint64_t *off, const uint32_t *col, const uint32_t *gid, uint32_t *idx,
                  uint32_t *pos, uint32_t *gidp, uint64_t *offp, uint32_t *colp, uint32_t *k0, uint32_t *v0,
                  uint32_t *k1, uint32_t *v1, uint64_t *sort_scratch, uint64_t *scan_scratch, hipStream_t s) {
  if (!n) return;
  const unsigned grid = grid_for(n) > 65536 ? 65536 : grid_for(n);
  hipLaunchKernelGGL(k_perm_sortkeys, dim3(grid), dim3(kBlock), 0, s, n, off, k0, v0);
  KC_LAUNCH_CHECK();
  const int which = radix_sort_pairs(n, 32, k0, v0, k1, v1, sort_scratch, s);
  const uint32_t *sorted = which ? v1 : v0;
  KC_HIP(hipMemcpyAsync(idx, sorted, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  uint32_t *degp = which ? k0 : k1;  // the other key buffer is free
  hipLaunchKernelGGL(k_perm_maps, dim3(grid), dim3(kBlock), 0, s, n, idx, gid, off, pos, gidp, degp);
  KC_LAUNCH_CHECK();
  uint64_t *total = scan_scratch;
  scan_exclusive_u32_to_u64(n, degp, offp, total, scan_scratch + 4, s);
  hipLaunchKernelGGL(k_copy_total, dim3(1), dim3(1), 0, s, total, offp + n);
  KC_LAUNCH_CHECK();
  const unsigned rows = n > (1u << 20) ? (1u << 20) : static_cast<unsigned>(n);
  hipLaunchKernelGGL(k_perm_rows, dim3(rows), dim3(64), 0, s, n, idx, off, col, pos, offp, colp);
  KC_LAUNCH_CHECK();
}

// Row ids of a CSR's slots, one wave per row.
__global__ __launch_bounds__(kBlock) void k_row_ids(uint64_t n, const uint64_t *off, uint32_t *row) {
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / kWave);
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kBlock / kWave) + threadIdx.x / kWave; v < n; v += waves)
    for (uint64_t e = off[v] + lane_id(); e < off[v + 1]; e += kWave) row[e] = static_cast<uint32_t>(v);
}

void sort_rows(uint64_t n, const uint64_t *off, uint32_t *col, uint64_t m2, uint32_t *k0, uint32_t *v0, uint32_t *k1,
               uint32_t *v1, uint64_t *scratch, hipStream_t s) {
  if (m2 < 2 || n < 2) return;
  int bits = 1;
  while (bits < 32 && (1ull << bits) < n) ++bits;
  const unsigned g = static_cast<unsigned>(std::min<uint64_t>(ceil_div(n, kBlock / kWave), 1u << 16));
  hipLaunchKernelGGL(k_row_ids, dim3(g), dim3(kBlock), 0, s, n, off, v0);
  KC_LAUNCH_CHECK();
  // (owner key, row) pairs by key, then stably by row: rows in order, each
  // grouped by owner part and ascending inside a group
  const int B = bits < 8 ? 8 : bits;
  const unsigned ge = static_cast<unsigned>(std::min<uint64_t>(grid_for(m2), 65536));
  hipLaunchKernelGGL(k_owner_keys, dim3(ge), dim3(kBlock), 0, s, m2, col, k0, B);
  KC_LAUNCH_CHECK();
  const int w = radix_sort_pairs(m2, B, k0, v0, k1, v1, scratch, s);
  uint32_t *sc = w ? k1 : k0, *sr = w ? v1 : v0, *tc = w ? k0 : k1, *tr = w ? v0 : v1;
  const int w2 = radix_sort_pairs(m2, bits, sr, sc, tr, tc, scratch, s);
  hipLaunchKernelGGL(k_owner_cols, dim3(ge), dim3(kBlock), 0, s, m2, w2 ? tc : sc, col, B);
  KC_LAUNCH_CHECK();
}

void identity_u32(uint64_t n, uint32_t *p, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_identity, dim3(grid_for(n) > 65536 ? 65536 : grid_for(n)), dim3(kBlock), 0, s, n, p);
  KC_LAUNCH_CHECK();
}

void labels_from_keys(uint64_t n, const uint64_t *key, uint32_t *labels, uint32_t *roots, uint32_t *partial,
                      hipStream_t s) {
  if (!n) return;
  const unsigned g = static_cast<unsigned>(std::min<uint64_t>(grid_for(n), kCountBlocks));
  hipLaunchKernelGGL(k_labels_from_keys, dim3(g), dim3(kBlock), 0, s, n, key, labels, partial);
  KC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kBlock), 0, s, partial, static_cast<uint64_t>(g), roots);
  KC_LAUNCH_CHECK();
}


uint64_t byte_positions(const uint8_t *b, uint64_t bytes, int what, uint64_t *pos, uint64_t *total, uint64_t *scratch,
                        hipStream_t s) {
  scan_exclusive<uint64_t>(bytes, ByteClassLoad{b, what}, PosStore{pos}, total, scratch, s);
  uint64_t h = 0;
  KC_HIP(hipMemcpyAsync(&h, total, sizeof(h), hipMemcpyDeviceToHost, s));
  KC_HIP(hipStreamSynchronize(s));
  return h;
}

void metis_vertex_lines(const uint8_t *b, uint64_t bytes, const uint64_t *nlpos, uint64_t lines, uint64_t n,
                        uint32_t *vline, uint64_t *total, uint64_t *scratch, hipStream_t s) {
  scan_exclusive<uint64_t>(lines, VertexLineLoad{b, nlpos, bytes}, VertexLineStore{vline, n}, total, scratch, s);
}

void metis_tokens(const uint8_t *b, uint64_t bytes, const uint64_t *tokpos, uint64_t count, const uint64_t *nlpos,
                  uint64_t nl, const uint32_t *vline, uint64_t n, uint32_t *tv, uint32_t *tval, int *bad,
                  hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_metis_tokens, dim3(static_cast<unsigned>(ceil_div(count, static_cast<uint64_t>(kBlock)))),
                     dim3(kBlock), 0, s, b, bytes, tokpos, count, nlpos, nl, vline, n, tv, tval, bad);
  KC_LAUNCH_CHECK();
}

void metis_other(const uint64_t *otherpos, uint64_t count, const uint64_t *nlpos, uint64_t nl, const uint32_t *vline,
                 int *bad, hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_metis_other, dim3(static_cast<unsigned>(ceil_div(count, static_cast<uint64_t>(kBlock)))),
                     dim3(kBlock), 0, s, otherpos, count, nlpos, nl, vline, bad);
  KC_LAUNCH_CHECK();
}

uint64_t metis_edges(uint64_t count, const uint32_t *tv, const uint32_t *tval, uint32_t *src, uint32_t *dst,
                     uint64_t *total, uint64_t *scratch, hipStream_t s) {
  scan_exclusive<uint64_t>(count, EdgeKeepLoad{tv}, EdgeStore{tv, tval, src, dst}, total, scratch, s);
  uint64_t h = 0;
  KC_HIP(hipMemcpyAsync(&h, total, sizeof(h), hipMemcpyDeviceToHost, s));
  KC_HIP(hipStreamSynchronize(s));
  return h;
}


uint64_t edge_text_lines(const uint8_t *b, uint64_t bytes, const uint64_t *nlpos, uint64_t nl, uint64_t lines,
                         uint64_t *eline, uint64_t *total, uint64_t *scratch, hipStream_t s) {
  scan_exclusive<uint64_t>(lines, EdgeLineLoad{EdgeLineSpan{b, nlpos, bytes, nl}}, PosStore{eline}, total, scratch,
                           s);
  uint64_t h = 0;
  KC_HIP(hipMemcpyAsync(&h, total, sizeof(h), hipMemcpyDeviceToHost, s));
  KC_HIP(hipStreamSynchronize(s));
  return h;
}

void edge_text(const uint8_t *b, uint64_t bytes, const uint64_t *nlpos, uint64_t nl, const uint64_t *eline,
               uint64_t count, uint64_t n, uint32_t *src, uint32_t *dst, int *bad, hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_edge_text, dim3(static_cast<unsigned>(ceil_div(count, static_cast<uint64_t>(kBlock)))),
                     dim3(kBlock), 0, s, EdgeLineSpan{b, nlpos, bytes, nl}, eline, count, n, src, dst, bad);
  KC_LAUNCH_CHECK();
}


// ---------------------------------------------------------------------------
// Hub replication (kc_options.replicate): a level-0 graph in which every hub
// is split into proxies that take over its edges.  Ids through gid (null:
// identity) feed the per-edge proxy pick, so the split does not depend on the
// storage order.

__global__ void k_find_hubs(uint64_t n, const uint64_t *off, uint64_t limit, uint32_t *list, uint32_t *count) {
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (v < n && off[v + 1] - off[v] >= limit) list[atomicAdd(count, 1u)] = static_cast<uint32_t>(v);
}

__global__ void k_hub_index(uint32_t H, const uint32_t *hubs, uint32_t *hubidx) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h < H) hubidx[hubs[h]] = h;
}

__device__ __forceinline__ uint32_t gid_of(const uint32_t *gid, uint32_t v) { return gid ? gid[v] : v; }

__global__ __launch_bounds__(kBlock) void k_replica_count(ReplicaArgs r, const HubEdgeChunk *chunks) {
  const HubEdgeChunk c = chunks[blockIdx.x];
  const uint32_t v = r.hubs[c.h], gv = gid_of(r.gid, v), k = r.k[c.h];
  const uint64_t base = r.slot_base[c.h];
  for (uint64_t e = c.e0 + threadIdx.x; e < c.e1; e += kBlock)
    atomicAdd(&r.cnt[base + replica_pick(gv, gid_of(r.gid, r.col[e]), r.seed, k)], 1u);
}

__global__ void k_replica_lengths(ReplicaArgs r, uint64_t n0, uint32_t R, uint32_t *len) {
  const uint64_t x = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (x < n0) {
    const uint32_t h = r.hubidx[x];
    len[x] = h == kNone ? static_cast<uint32_t>(r.off[x + 1] - r.off[x]) : r.nexist[h];
  } else if (x < n0 + R) {
    len[x] = 1 + r.cnt[r.slot_of_rep[x - n0]];
  }
}

// The proxy of hub w that holds w's edge to u.
__device__ __forceinline__ uint32_t proxy_of(const ReplicaArgs &r, uint32_t w, uint32_t hw, uint32_t gu) {
  return r.rep_of_slot[r.slot_base[hw] + replica_pick(gid_of(r.gid, w), gu, r.seed, r.k[hw])];
}

// Rows of the other vertices, one wave per row: the neighbours in place, a hub
// replaced by its proxy for this edge.
__global__ __launch_bounds__(kBlock) void k_replica_rows(ReplicaArgs r, uint64_t n0, const uint64_t *off_r,
                                                          uint32_t *col_r) {
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / kWave);
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kBlock / kWave) + threadIdx.x / kWave; v < n0; v += waves) {
    if (r.hubidx[v] != kNone) continue;
    const uint64_t e0 = r.off[v], e1 = r.off[v + 1], o = off_r[v];
    const uint32_t gv = gid_of(r.gid, static_cast<uint32_t>(v));
    for (uint64_t e = e0 + lane_id(); e < e1; e += kWave) {
      const uint32_t w = r.col[e], hw = r.hubidx[w];
      col_r[o + (e - e0)] = hw == kNone ? w : proxy_of(r, w, hw, gv);
    }
  }
}

// Hub rows list their proxies; a proxy row starts with its hub; ids of the
// level (a proxy takes its hub's).
__global__ void k_replica_fixed(ReplicaArgs r, uint64_t n0, uint32_t H, uint32_t R, const uint64_t *off_r,
                                uint32_t *col_r, uint32_t *gid_r) {
  const uint64_t x = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (x < n0) gid_r[x] = gid_of(r.gid, static_cast<uint32_t>(x));
  if (x < R) {
    const uint32_t root = r.root_of_rep[x];
    col_r[off_r[n0 + x]] = root;
    gid_r[n0 + x] = gid_of(r.gid, root);
  }
  if (x < H) {
    const uint64_t o = off_r[r.hubs[x]];
    for (uint32_t j = 0; j < r.nexist[x]; ++j) col_r[o + j] = r.first_rep[x] + j;
  }
}

// The hub edges, each appended to the row of its proxy (after the hub entry).
__global__ __launch_bounds__(kBlock) void k_replica_edges(ReplicaArgs r, const HubEdgeChunk *chunks,
                                                          const uint64_t *off_r, uint32_t *cursor, uint32_t *col_r) {
  const HubEdgeChunk c = chunks[blockIdx.x];
  const uint32_t v = r.hubs[c.h], gv = gid_of(r.gid, v), k = r.k[c.h];
  const uint64_t base = r.slot_base[c.h];
  for (uint64_t e = c.e0 + threadIdx.x; e < c.e1; e += kBlock) {
    const uint32_t w = r.col[e], hw = r.hubidx[w];
    const uint64_t slot = base + replica_pick(gv, gid_of(r.gid, w), r.seed, k);
    const uint32_t p = r.rep_of_slot[slot];
    col_r[off_r[p] + 1 + atomicAdd(&cursor[slot], 1u)] = hw == kNone ? w : proxy_of(r, w, hw, gv);
  }
}

__global__ void k_row_bounds(uint32_t count, const uint32_t *rows, const uint64_t *off, uint64_t *bounds) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    bounds[2ull * i] = off[rows[i]];
    bounds[2ull * i + 1] = off[rows[i] + 1];
  }
}

void row_bounds(uint32_t count, const uint32_t *rows, const uint64_t *off, uint64_t *bounds, hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_row_bounds, dim3(static_cast<unsigned>(ceil_div(count, kBlock))), dim3(kBlock), 0, s, count,
                     rows, off, bounds);
  KC_LAUNCH_CHECK();
}

void find_hubs(uint64_t n, const uint64_t *off, uint64_t limit, uint32_t *list, uint32_t *count, hipStream_t s) {
  fill_u32(count, 1, 0, s);
  if (!n) return;
  hipLaunchKernelGGL(k_find_hubs, dim3(grid_for(n)), dim3(kBlock), 0, s, n, off, limit, list, count);
  KC_LAUNCH_CHECK();
}

void replica_count(const ReplicaArgs &r, uint64_t n0, uint32_t H, const HubEdgeChunk *chunks, uint32_t n_chunks,
                   hipStream_t s) {
  fill_u32(r.hubidx, n0, kNone, s);
  if (H) {
    hipLaunchKernelGGL(k_hub_index, dim3(static_cast<unsigned>(ceil_div(H, kBlock))), dim3(kBlock), 0, s, H, r.hubs,
                       r.hubidx);
    KC_LAUNCH_CHECK();
  }
  if (n_chunks) {
    hipLaunchKernelGGL(k_replica_count, dim3(n_chunks), dim3(kBlock), 0, s, r, chunks);
    KC_LAUNCH_CHECK();
  }
}

void replica_lengths(const ReplicaArgs &r, uint64_t n0, uint32_t R, uint32_t *len, hipStream_t s) {
  if (!(n0 + R)) return;
  hipLaunchKernelGGL(k_replica_lengths, dim3(grid_for(n0 + R)), dim3(kBlock), 0, s, r, n0, R, len);
  KC_LAUNCH_CHECK();
}

void replica_fill(const ReplicaArgs &r, uint64_t n0, uint32_t H, uint32_t R, const HubEdgeChunk *chunks,
                  uint32_t n_chunks, const uint64_t *off_r, uint32_t *cursor, uint32_t *col_r, uint32_t *gid_r,
                  hipStream_t s) {
  if (n0) {
    const unsigned g = static_cast<unsigned>(std::min<uint64_t>(ceil_div(n0, kBlock / kWave), 1u << 16));
    hipLaunchKernelGGL(k_replica_rows, dim3(g), dim3(kBlock), 0, s, r, n0, off_r, col_r);
    KC_LAUNCH_CHECK();
  }
  const uint64_t m = std::max<uint64_t>({n0, static_cast<uint64_t>(H), static_cast<uint64_t>(R)});
  if (m) {
    hipLaunchKernelGGL(k_replica_fixed, dim3(grid_for(m)), dim3(kBlock), 0, s, r, n0, H, R, off_r, col_r, gid_r);
    KC_LAUNCH_CHECK();
  }
  if (n_chunks) {
    hipLaunchKernelGGL(k_replica_edges, dim3(n_chunks), dim3(kBlock), 0, s, r, chunks, off_r, cursor, col_r);
    KC_LAUNCH_CHECK();
  }
}

}  // namespace kc
