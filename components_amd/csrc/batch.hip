// Batched exponential contraction: thousands of independent graphs packed as
// CSR segments, contracted in one persistent launch.
//
// Every workgroup takes graphs from a queue (largest first) and runs the whole
// level loop of exponential contraction on one graph with block barriers
// between the phases: keys from the deviates, in-place sweeps to the fixed
// point, cluster labels, cluster edges deduplicated in a per-graph hash table,
// the next level's CSR, and at the end the undo and the canonical labels.  The
// per-graph decisions are the plain path's (same deviates with the graph's
// local ids, same tie-break by level index, same stall rule), so each graph's
// levels and labels equal what the plain path gives on that graph alone with
// no base case.  No workgroup waits on another, so every launch shape drains.
//
// Reference counterpart per graph: ExponentialContraction::FindComponents
// (lib/components/exponential_contraction.h:60-319) with RunContractionLP
// (:439-628), DynamicContraction::ExponentialContraction and UndoContraction
// (lib/contraction/dynamic_contraction.h:62-136, :568-658).  The batch itself
// has no reference counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "csr_build.h"
#include "deviates.h"
#include "kc_common.h"
#include "kernels.h"

namespace kc {

// Per-call parameters, read from device memory so that a captured graph can
// be replayed with new seeds.
struct BatchParams {
  int32_t seed;
  int32_t fused;
  int32_t stream;
  uint32_t k_max;
  const double *table;
  uint32_t max_levels;
};

struct BatchDev {
  uint32_t graphs;
  uint32_t queued;        // graphs the persistent kernel takes (the first entries of order)
  const uint32_t *order;  // those graphs by decreasing edge count
  const uint64_t *vtx_off, *edge_off;
  const uint64_t *off;  // global row offsets into col (sum n + 1)
  const uint32_t *col;  // local neighbour ids
  const uint32_t *nonisolated;
  uint32_t *labels;  // sum n: local canonical labels
  uint64_t *key;     // sum n (rep arrays after the level loop)
  uint32_t *loff[2];  // sum (n + 1) each: level CSR offsets
  uint32_t *lcol[2];  // sum 2m each: level CSR columns
  uint32_t *cnt;      // sum n: unique cluster edges per centre, then cursors, then min ids
  uint64_t *table;    // per graph at table_off[g]: cluster-edge hash set (all empty between calls)
  const uint64_t *table_off;
  uint32_t *stack;  // per graph at 10 * vtx_off[g]: level records for the undo
  uint32_t *glevels, *gsweeps, *gradius, *gstatus, *gcomp;
  uint32_t *queue;
  const BatchParams *params;
};

namespace {

constexpr uint64_t kEmpty = ~0ull;
constexpr uint32_t kStackPerVertex = 10;
constexpr uint32_t kStatusOverflow = 1, kStatusLevels = 2, kStatusTable = 3;

__device__ __forceinline__ uint64_t pair_hash(uint64_t k) {
  k ^= k >> 31;
  k *= 0x7fb5d329728ea185ull;
  k ^= k >> 27;
  k *= 0x81dadef4bc2dd44dull;
  return k ^ (k >> 33);
}

// Inserts k: 1 when this call added it, 0 when it was present, 2 when every
// slot was probed (cannot happen below load 1/2; reported as an error).
// Reads before every CAS.
__device__ __forceinline__ uint32_t table_insert(uint64_t *tab, uint32_t mask, uint64_t k) {
  uint32_t h = static_cast<uint32_t>(pair_hash(k)) & mask;
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    uint64_t cur = tab[h];
    if (cur == k) return 0;
    if (cur == kEmpty) {
      cur = atomicCAS(reinterpret_cast<unsigned long long *>(&tab[h]), static_cast<unsigned long long>(kEmpty),
                      static_cast<unsigned long long>(k));
      if (cur == kEmpty) return 1;
      if (cur == k) return 0;
    }
    h = (h + 1) & mask;
  }
  return 2;
}

// Block-wide exclusive scan of one u64 per thread, with a running carry in
// LDS.  All threads of the block call it with the same trip count.
template <int B>
__device__ __forceinline__ uint64_t block_scan(uint64_t x, uint64_t *s_wave, uint64_t &s_carry, uint64_t &total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t incl = x;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint64_t before = s_carry, sum = 0;
  for (int w = 0; w < B / 64; ++w) {
    const uint64_t t = s_wave[w];
    if (w < wave) before += t;
    sum += t;
  }
  total = sum;
  __syncthreads();
  if (threadIdx.x == 0) s_carry += sum;
  return before + incl - x;
}

struct Shared {
  uint64_t wave[16];
  uint64_t carry;
  uint32_t graph;
  uint32_t flag;
  uint32_t radius;
  uint32_t full;
  uint32_t roots;
};

// One in-place sweep over the level (one lane per row); returns through
// s.flag whether a key changed.
template <int B>
__device__ void sweep_level(uint32_t nl, const uint32_t *off, const uint32_t *col, uint64_t *key, Shared &s) {
  for (uint32_t v = threadIdx.x; v < nl; v += B) {
    const uint32_t e0 = off[v], e1 = off[v + 1];
    if (e1 == e0) continue;
    const uint64_t own = key[v];
    uint64_t best = own;
    uint32_t e = e0;
    for (; e + 4 <= e1; e += 4) {  // four independent gathers in flight
      const uint32_t c0 = col[e], c1 = col[e + 1], c2 = col[e + 2], c3 = col[e + 3];
      const uint64_t k0 = key[c0], k1 = key[c1], k2 = key[c2], k3 = key[c3];
      const uint64_t a = (k0 < k1 ? k0 : k1), b = (k2 < k3 ? k2 : k3);
      const uint64_t c = (a < b ? a : b) + kOneHop;
      best = c < best ? c : best;
    }
    for (; e < e1; ++e) {
      const uint64_t c = key[col[e]] + kOneHop;
      best = c < best ? c : best;
    }
    if (best < own) {
      key[v] = best;
      s.flag = 1;
    }
  }
}

// Cluster edges of the level into the hash set (one lane per row); cnt[c]
// counts the distinct cluster edges leaving centre c.
template <int B>
__device__ void insert_level(uint32_t nl, const uint32_t *off, const uint32_t *col, const uint32_t *lab,
                             uint64_t *tab, uint32_t mask, uint32_t *cnt, Shared &s) {
  for (uint32_t v = threadIdx.x; v < nl; v += B) {
    const uint32_t e0 = off[v], e1 = off[v + 1];
    if (e1 == e0) continue;
    const uint32_t lv = lab[v];
    uint32_t added = 0;
    for (uint32_t e = e0; e < e1; ++e) {
      const uint32_t lw = lab[col[e]];
      if (lw == lv) continue;
      const uint32_t r = table_insert(tab, mask, (static_cast<uint64_t>(lv) << 32) | lw);
      added += r == 1;
      if (r == 2) s.full = 1;
    }
    if (added) atomicAdd(&cnt[lv], added);
  }
}

template <int B>
__device__ void contract_graph(const BatchDev &d, const BatchParams &p, uint32_t g, Shared &s) {
  const uint32_t tid = threadIdx.x;
  const uint64_t v0 = d.vtx_off[g];
  const uint32_t n0 = static_cast<uint32_t>(d.vtx_off[g + 1] - v0);
  const uint64_t eb = d.edge_off[g];
  if (n0 == 0) {
    if (tid == 0) {
      d.glevels[g] = 0;
      d.gsweeps[g] = 0;
      d.gradius[g] = 0;
      d.gstatus[g] = 0;
      d.gcomp[g] = 0;
    }
    return;
  }
  uint64_t *key = d.key + v0;
  uint32_t *cnt = d.cnt + v0;
  uint32_t *stack = d.stack + kStackPerVertex * v0;
  const uint64_t cap = static_cast<uint64_t>(kStackPerVertex) * n0;
  uint64_t *tab = d.table + d.table_off[g];
  const uint64_t tcap = d.table_off[g + 1] - d.table_off[g];
  uint32_t *offb[2] = {d.loff[0] + v0 + g, d.loff[1] + v0 + g};
  uint32_t *colb[2] = {d.lcol[0] + eb, d.lcol[1] + eb};
  const uint64_t rng_offset = d.nonisolated[g];
  uint32_t *labels = d.labels + v0;
  if (rng_offset == 0) {  // no edge: every vertex is its own component, no level runs
    for (uint32_t v = tid; v < n0; v += B) labels[v] = v;
    if (tid == 0) {
      d.glevels[g] = 0;
      d.gsweeps[g] = 0;
      d.gradius[g] = 0;
      d.gstatus[g] = 0;
      d.gcomp[g] = n0;
    }
    __syncthreads();
    return;
  }

  // level 0: the input rows with local offsets
  for (uint32_t v = tid; v <= n0; v += B) offb[0][v] = static_cast<uint32_t>(d.off[v0 + v] - eb);
  __syncthreads();

  uint32_t nl = n0;
  const uint32_t *off = offb[0];
  const uint32_t *col = d.col + eb;
  uint32_t depth = 0;  // contractions done: level CSR offsets in offb[depth & 1], columns in colb[(depth - 1) & 1]
  uint64_t rec = 0, last = kNone;
  uint32_t iteration = 1, levels = 0, sweeps = 0, radius_sum = 0, status = 0;
  bool level0 = true;
  while (true) {
    if (levels >= p.max_levels) {
      status = kStatusLevels;
      break;
    }
    const uint64_t hdr = level0 ? 2 : 2 + static_cast<uint64_t>(nl);
    if (rec + hdr + 2ull * nl > cap) {
      status = kStatusOverflow;
      break;
    }
    const uint32_t *gid = level0 ? nullptr : stack + rec + 2;
    uint32_t *lab = stack + rec + hdr;
    uint32_t *newidx = lab + nl;
    for (uint32_t v = tid; v < nl; v += B) {
      const uint32_t id = gid ? gid[v] : v;
      const uint32_t dev = exp_deviate(id, p.seed, iteration, rng_offset, p.fused != 0, p.stream, p.table, p.k_max);
      key[v] = (static_cast<uint64_t>(dev) << 32) | v;
    }
    if (tid == 0) {
      s.radius = 0;
      s.full = 0;
    }
    __syncthreads();
    uint32_t sw = 0;
    bool first_change = false;
    while (true) {
      if (tid == 0) s.flag = 0;
      __syncthreads();
      sweep_level<B>(nl, off, col, key, s);
      __syncthreads();
      const bool changed = s.flag != 0;
      __syncthreads();
      ++sw;
      if (sw == 1) first_change = changed;
      if (!changed) break;
    }
    sweeps += sw;
    ++levels;
    if (!first_change && (!level0 || rng_offset == n0)) {
      // stall: every vertex is its own centre and keeps an edge, so the
      // contraction is the identity; draw again on the same graph
      ++iteration;
      continue;
    }
    // labels and the radius; counters cleared
    for (uint32_t v = tid; v < nl; v += B) {
      const uint64_t k = key[v];
      const uint32_t c = static_cast<uint32_t>(k);
      lab[v] = c;
      const uint32_t r = static_cast<uint32_t>(k >> 32) - static_cast<uint32_t>(key[c] >> 32);
      if (r) atomicMax(&s.radius, r);
      cnt[v] = 0;
    }
    __syncthreads();
    radius_sum += s.radius;
    const uint32_t m2l = off[nl];
    uint64_t tsz = 64;
    while (tsz < 2ull * m2l) tsz <<= 1;
    if (tsz > tcap) tsz = tcap;
    insert_level<B>(nl, off, col, lab, tab, static_cast<uint32_t>(tsz - 1), cnt, s);
    __syncthreads();
    if (s.full) {  // the host clears the hash sets after a failed run
      status = kStatusTable;
      break;
    }
    // next-level ids (centres with a cluster edge, in label order), offsets, gids
    const uint64_t rec_next = rec + hdr + 2ull * nl;
    const bool room = rec_next + 2 + 3ull * nl <= cap;
    uint32_t *off_next = offb[(depth + 1) & 1];
    uint32_t *col_next = colb[depth & 1];
    uint32_t *gid_next = stack + rec_next + 2;
    if (tid == 0) s.carry = 0;
    __syncthreads();
    uint64_t total = 0;
    for (uint32_t base = 0; base < nl; base += B) {
      const uint32_t c = base + tid;
      const uint32_t k = c < nl ? cnt[c] : 0u;
      const uint64_t x = k ? ((1ull << 32) | k) : 0ull;
      const uint64_t ex = block_scan<B>(x, s.wave, s.carry, total);
      if (c < nl) {
        const uint32_t ni = static_cast<uint32_t>(ex >> 32);
        newidx[c] = k ? ni : kNone;
        if (k) {
          off_next[ni] = static_cast<uint32_t>(ex);
          cnt[c] = static_cast<uint32_t>(ex);
          if (room) gid_next[ni] = gid ? gid[c] : c;
        }
      }
    }
    __syncthreads();
    const uint64_t tot = s.carry;
    const uint32_t n_next = static_cast<uint32_t>(tot >> 32), m2_next = static_cast<uint32_t>(tot);
    if (tid == 0) {
      off_next[n_next] = m2_next;
      stack[rec] = nl;
      stack[rec + 1] = static_cast<uint32_t>(last);
    }
    // next-level columns; the hash set is left empty
    for (uint64_t slot = tid; slot < tsz; slot += B) {
      const uint64_t t = tab[slot];
      if (t != kEmpty) {
        const uint32_t lu = static_cast<uint32_t>(t >> 32), lw = static_cast<uint32_t>(t);
        const uint32_t pos = atomicAdd(&cnt[lu], 1u);
        col_next[pos] = newidx[lw];
        tab[slot] = kEmpty;
      }
    }
    __syncthreads();
    last = rec;
    if (n_next == 0) break;
    if (!room) {
      status = kStatusOverflow;
      break;
    }
    rec = rec_next;
    nl = n_next;
    off = off_next;
    col = col_next;
    ++depth;
    level0 = false;
    ++iteration;
  }
  // undo, top down: rep_l[v] = rep_{l+1}[newidx[lab[v]]], or gid[lab[v]]
  uint32_t *repa = reinterpret_cast<uint32_t *>(key);
  uint32_t *repb = repa + n0;
  uint32_t *up = nullptr;
  uint64_t r = status ? kNone : last;
  while (r != kNone) {
    const uint32_t n = stack[r];
    const uint64_t prev = stack[r + 1] == kNone ? kNone : stack[r + 1];
    const bool lvl0 = r == 0;
    const uint32_t *gid = lvl0 ? nullptr : stack + r + 2;
    const uint32_t *lab = stack + r + (lvl0 ? 2 : 2 + static_cast<uint64_t>(n));
    const uint32_t *newidx = lab + n;
    uint32_t *out = up == repa ? repb : repa;
    for (uint32_t v = tid; v < n; v += B) {
      const uint32_t a = lab[v], a2 = newidx[a];
      out[v] = a2 != kNone ? up[a2] : (gid ? gid[a] : a);
    }
    __syncthreads();
    up = out;
    r = prev;
  }
  if (!up) {
    for (uint32_t v = tid; v < n0; v += B) labels[v] = kNone;
  } else {
    // canonical: minimum local id per representative
    for (uint32_t v = tid; v < n0; v += B) cnt[v] = kNone;
    if (tid == 0) s.roots = 0;
    __syncthreads();
    for (uint32_t v = tid; v < n0; v += B) atomicMin(&cnt[up[v]], v);
    __syncthreads();
    uint32_t roots = 0;
    for (uint32_t v = tid; v < n0; v += B) {
      const uint32_t l = cnt[up[v]];
      labels[v] = l;
      roots += l == v;
    }
    if (roots) atomicAdd(&s.roots, roots);
    __syncthreads();
  }
  if (tid == 0) {
    d.gcomp[g] = up ? s.roots : 0;
    d.glevels[g] = levels;
    d.gsweeps[g] = sweeps;
    d.gradius[g] = radius_sum;
    d.gstatus[g] = status;
  }
  __syncthreads();
}

template <int B>
__global__ __launch_bounds__(B) void k_batch(BatchDev d) {
  __shared__ Shared s;
  const BatchParams p = *d.params;
  while (true) {
    if (threadIdx.x == 0) s.graph = atomicAdd(d.queue, 1u);
    __syncthreads();
    const uint32_t qi = s.graph;
    __syncthreads();
    if (qi >= d.queued) break;
    contract_graph<B>(d, p, d.order[qi], s);
  }
}

// Clears the hash sets of graphs that stopped early (their tables may hold
// entries) -- launched only after a run reported a non-zero status.
__global__ void k_table_clear(uint64_t n, uint64_t *tab) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    tab[i] = kEmpty;
}

// Union of the large graphs: union vertex u (or edge j) lies in large graph k,
// the largest k with uoff[k] <= u (ueoff[k] <= j); its rows and columns are
// those of batch graph big[k], shifted to the union's numbering.
__device__ __forceinline__ uint32_t union_part(const uint64_t *start, uint32_t parts, uint64_t x) {
  uint32_t lo = 0, hi = parts;  // start[lo] <= x < start[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (start[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// The same for the lanes of a wave holding consecutive x from `first` (the
// wave's first lane) on: one search for `first`, then each lane steps forward
// over the parts it passes (the large graphs exceed the split size, so a wave
// spans one or two of them).
__device__ __forceinline__ uint32_t union_part_wave(const uint64_t *start, uint32_t parts, uint64_t x, uint64_t first) {
  uint32_t k = union_part(start, parts, first);
  while (k + 1 < parts && start[k + 1] <= x) ++k;
  return k;
}

__global__ void k_union_rows(uint64_t n_union, uint32_t parts, const uint64_t *uoff, const uint64_t *ueoff,
                             const uint32_t *big, const uint64_t *vtx_off, const uint64_t *edge_off,
                             const uint64_t *off, uint64_t *uoff_rows) {
  for (uint64_t u = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; u <= n_union;
       u += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (u == n_union) {
      uoff_rows[u] = ueoff[parts];
      continue;
    }
    const uint32_t k = union_part_wave(uoff, parts, u, u - lane_id());
    const uint32_t g = big[k];
    uoff_rows[u] = off[vtx_off[g] + (u - uoff[k])] - edge_off[g] + ueoff[k];
  }
}

__global__ void k_union_cols(uint64_t m2_union, uint32_t parts, const uint64_t *uoff, const uint64_t *ueoff,
                             const uint32_t *big, const uint64_t *edge_off, const uint32_t *col, uint32_t *ucol) {
  for (uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; j < m2_union;
       j += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t k = union_part_wave(ueoff, parts, j, j - lane_id());
    ucol[j] = col[edge_off[big[k]] + (j - ueoff[k])] + static_cast<uint32_t>(uoff[k]);
  }
}

// Statistics of the large graphs: the union run's levels, sweeps and radius
// sum, status 0, and the component count zeroed for k_union_labels.
__global__ void k_union_stats(uint32_t parts, const uint32_t *big, uint32_t levels, uint32_t sweeps, uint32_t radius,
                              uint32_t *glevels, uint32_t *gsweeps, uint32_t *gradius, uint32_t *gstatus,
                              uint32_t *gcomp) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= parts) return;
  const uint32_t g = big[k];
  glevels[g] = levels;
  gsweeps[g] = sweeps;
  gradius[g] = radius;
  gstatus[g] = 0;
  gcomp[g] = 0;
}

// Local canonical labels of the large graphs from the union's (a component
// never leaves its graph, so its minimum union id lies in the same graph),
// and each graph's component count (vertices that are their own label).
__global__ void k_union_labels(uint64_t n_union, uint32_t parts, const uint64_t *uoff, const uint32_t *big,
                               const uint64_t *vtx_off, const uint32_t *ulab, uint32_t *labels, uint32_t *gcomp) {
  const uint64_t u = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  const bool act = u < n_union;
  uint32_t g = 0;
  bool root = false;
  if (act) {
    const uint32_t k = union_part_wave(uoff, parts, u, u - lane_id());
    g = big[k];
    const uint32_t l = ulab[u];
    labels[vtx_off[g] + (u - uoff[k])] = l - static_cast<uint32_t>(uoff[k]);
    root = l == static_cast<uint32_t>(u);
  }
  // one atomic per wave when the wave's vertices share a graph
  const uint32_t g0 = __shfl(g, 0);
  const bool same = __ballot(act && g != g0) == 0;
  if (same) {
    const uint32_t c = static_cast<uint32_t>(__popcll(__ballot(root)));
    if (lane_id() == 0 && c) atomicAdd(&gcomp[g0], c);
  } else if (root) {
    atomicAdd(&gcomp[g], 1u);
  }
}

}  // namespace

void batch_launch(int block, uint32_t grid, const BatchDev &d, hipStream_t s) {
  switch (block) {
    case 256:
      hipLaunchKernelGGL(k_batch<256>, dim3(grid), dim3(256), 0, s, d);
      break;
    case 512:
      hipLaunchKernelGGL(k_batch<512>, dim3(grid), dim3(512), 0, s, d);
      break;
    case 1024:
      hipLaunchKernelGGL(k_batch<1024>, dim3(grid), dim3(1024), 0, s, d);
      break;
    default:
      fail(KC_ERR_INVALID, "batch: unsupported block size %d", block);
  }
  KC_LAUNCH_CHECK();
}

void batch_table_clear(uint64_t n, uint64_t *tab, hipStream_t s) {
  if (!n) return;
  const uint64_t g = std::min<uint64_t>(ceil_div(n, kBlock), 65536);
  hipLaunchKernelGGL(k_table_clear, dim3(static_cast<unsigned>(g)), dim3(kBlock), 0, s, n, tab);
  KC_LAUNCH_CHECK();
}

}  // namespace kc

// ===========================================================================
// Host side.
using namespace kc;

struct kc_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t graphs = 0;
  uint64_t n = 0, m2 = 0;
  int cus = 0;
  std::vector<uint64_t> h_vtx_off, h_edge_off;
  std::vector<void *> owned;
  uint32_t *order_d = nullptr;
  // graphs above the split size go through the device-wide path as one union
  bool split_valid = false;
  uint32_t split_built = 0;
  kc_graph *big = nullptr;
  uint32_t n_big = 0;
  uint64_t n_union = 0, m2_union = 0;
  uint32_t *d_big = nullptr;
  uint64_t *d_uoff = nullptr, *d_ueoff = nullptr;
  BatchDev dev{};
  BatchParams *d_params = nullptr, *h_params = nullptr;
  uint64_t *table = nullptr;
  uint64_t table_words = 0;
  uint32_t *queue = nullptr;
  double *d_table = nullptr;
  uint64_t table_cap = 0;
  double table_beta = -1.0;
  uint32_t table_k = 0;
  kc_batch_tuning tuning{256, 4, 1, 0};
  hipGraphExec_t exec = nullptr;
  kc_batch_tuning exec_tuning{0, 0, 0, 0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool ran = false;

  template <class T>
  T *dalloc(uint64_t count) {
    void *p = nullptr;
    const size_t bytes = std::max<uint64_t>(count, 1) * sizeof(T) + 64;
    if (hipMalloc(&p, bytes) != hipSuccess) fail(KC_ERR_OOM, "batch: hipMalloc of %zu bytes failed", bytes);
    owned.push_back(p);
    return static_cast<T *>(p);
  }
  ~kc_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (big) (void)kc_graph_destroy(big);
    if (exec) (void)hipGraphExecDestroy(exec);
    for (void *p : owned) (void)hipFree(p);
    if (h_params) (void)hipHostFree(h_params);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

void check_tuning(const kc_batch_tuning &t) {
  if (t.block != 256 && t.block != 512 && t.block != 1024)
    fail(KC_ERR_INVALID, "batch: block must be 256, 512 or 1024 (got %d)", t.block);
  if (t.groups_per_cu < 1 || t.groups_per_cu > 16)
    fail(KC_ERR_INVALID, "batch: groups_per_cu must be in [1, 16] (got %d)", t.groups_per_cu);
}

void batch_params(kc_batch *b, const kc_options &o) {
  if (o.variant != KC_VARIANT_EXP) fail(KC_ERR_UNSUPPORTED, "batch: only exponential contraction is batched");
  if (o.learned) fail(KC_ERR_UNSUPPORTED, "batch: learned centre proposals are not batched");
  if (b->table_beta != o.beta || !b->d_table) {
    if (!(o.beta > 1e-6) || !(o.beta < 1e6)) fail(KC_ERR_INVALID, "beta must be in (1e-6, 1e6), got %g", o.beta);
    const std::vector<double> t = build_deviate_table(o.beta);
    KC_HIP(hipStreamSynchronize(b->stream));  // the old table may still be read
    if (t.size() + 1 > b->table_cap) {
      b->table_cap = t.size() + 1;
      b->d_table = b->dalloc<double>(b->table_cap);
    }
    if (!t.empty()) KC_HIP(hipMemcpy(b->d_table, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    b->table_k = static_cast<uint32_t>(t.size());
    b->table_beta = o.beta;
  }
  KC_HIP(hipStreamSynchronize(b->stream));  // h_params may still be in flight
  BatchParams &p = *b->h_params;
  p.seed = o.seed;
  p.fused = o.deviate_fma;
  p.stream = o.deviate_stream;
  p.k_max = b->table_k;
  p.table = b->d_table;
  p.max_levels = o.max_levels > 0 ? static_cast<uint32_t>(o.max_levels) : 4096u;
  KC_HIP(hipMemcpyAsync(b->d_params, b->h_params, sizeof(BatchParams), hipMemcpyHostToDevice, b->stream));
}

// Routes the graphs with more than `split` vertices (none when 0) to the
// device-wide path: the persistent kernel's queue gets the others, largest
// first, and the large ones are copied into one union handle.
void prepare_split(kc_batch *b, uint32_t split) {
  if (b->split_valid && b->split_built == split) return;
  b->split_valid = false;
  std::vector<uint32_t> small, bigs;
  for (uint32_t g = 0; g < b->graphs; ++g) {
    const uint64_t ng = b->h_vtx_off[g + 1] - b->h_vtx_off[g];
    (split && ng > split ? bigs : small).push_back(g);
  }
  std::stable_sort(small.begin(), small.end(), [&](uint32_t x, uint32_t y) {
    return b->h_edge_off[x + 1] - b->h_edge_off[x] > b->h_edge_off[y + 1] - b->h_edge_off[y];
  });
  KC_HIP(hipStreamSynchronize(b->stream));  // the old queue order may still be read
  if (!small.empty()) KC_HIP(hipMemcpy(b->order_d, small.data(), 4ull * small.size(), hipMemcpyHostToDevice));
  b->dev.queued = static_cast<uint32_t>(small.size());
  if (b->big) {
    (void)kc_graph_destroy(b->big);
    b->big = nullptr;
  }
  b->n_big = static_cast<uint32_t>(bigs.size());
  b->n_union = b->m2_union = 0;
  if (b->n_big) {
    std::vector<uint64_t> uoff(b->n_big + 1, 0), ueoff(b->n_big + 1, 0);
    for (uint32_t k = 0; k < b->n_big; ++k) {
      const uint32_t g = bigs[k];
      uoff[k + 1] = uoff[k] + (b->h_vtx_off[g + 1] - b->h_vtx_off[g]);
      ueoff[k + 1] = ueoff[k] + (b->h_edge_off[g + 1] - b->h_edge_off[g]);
    }
    b->n_union = uoff[b->n_big];
    b->m2_union = ueoff[b->n_big];
    KC_HIP(hipMemcpy(b->d_big, bigs.data(), 4ull * b->n_big, hipMemcpyHostToDevice));
    KC_HIP(hipMemcpy(b->d_uoff, uoff.data(), 8ull * (b->n_big + 1), hipMemcpyHostToDevice));
    KC_HIP(hipMemcpy(b->d_ueoff, ueoff.data(), 8ull * (b->n_big + 1), hipMemcpyHostToDevice));
    uint64_t *off = nullptr;
    uint32_t *col = nullptr;
    hipStream_t gs = nullptr;
    b->big = graph_device_begin(b->n_union, b->m2_union, &off, &col, &gs);
    const uint32_t gr = static_cast<uint32_t>(std::min<uint64_t>(ceil_div(b->n_union + 1, kBlock), 65536));
    hipLaunchKernelGGL(k_union_rows, dim3(gr), dim3(kBlock), 0, gs, b->n_union, b->n_big, b->d_uoff, b->d_ueoff,
                       b->d_big, b->dev.vtx_off, b->dev.edge_off, b->dev.off, off);
    KC_LAUNCH_CHECK();
    if (b->m2_union) {
      const uint32_t gc = static_cast<uint32_t>(std::min<uint64_t>(ceil_div(b->m2_union, kBlock), 65536));
      hipLaunchKernelGGL(k_union_cols, dim3(gc), dim3(kBlock), 0, gs, b->m2_union, b->n_big, b->d_uoff, b->d_ueoff,
                         b->d_big, b->dev.edge_off, b->dev.col, col);
      KC_LAUNCH_CHECK();
    }
    graph_device_finish(b->big);
  }
  b->split_built = split;
  b->split_valid = true;
}

void enqueue(kc_batch *b, const kc_batch_tuning &t) {
  fill_u32(b->queue, 1, 0, b->stream);
  const uint32_t grid = static_cast<uint32_t>(b->cus) * static_cast<uint32_t>(t.groups_per_cu);
  batch_launch(t.block, grid, b->dev, b->stream);
}

// One run on the current parameters: eager launches, or the captured graph
// (captured again when the launch shape changed).
void launch(kc_batch *b) {
  const kc_batch_tuning &t = b->tuning;
  if (!t.use_graph) {
    enqueue(b, t);
    return;
  }
  if (!b->exec || std::memcmp(&b->exec_tuning, &t, sizeof(t)) != 0) {
    if (b->exec) {
      KC_HIP(hipGraphExecDestroy(b->exec));
      b->exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    KC_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue(b, t);
    } catch (...) {
      (void)hipStreamEndCapture(b->stream, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    KC_HIP(hipStreamEndCapture(b->stream, &graph));
    const hipError_t e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) fail(KC_ERR_HIP, "batch: hipGraphInstantiate failed: %s", hipGetErrorString(e));
    b->exec_tuning = t;
  }
  KC_HIP(hipGraphLaunch(b->exec, b->stream));
}

// Runs once and checks every graph's status; returns the device time.  The
// persistent kernel takes the small graphs while the device-wide path runs the
// union of the large ones on the union handle's stream.
float run_once(kc_batch *b, const kc_options &o) {
  prepare_split(b, b->tuning.split_vertices);
  KC_HIP(hipEventRecord(b->ev0, b->stream));
  launch(b);
  if (b->big) {
    kc_options ob = o;
    ob.use_bfs = 1;
    ob.record_trace = 0;
    ob.time_kernels = 0;
    ob.learned = 0;
    ob.seq_limit = std::max<uint64_t>(o.seq_limit, b->m2_union / 2 / 256);
    kc_stats st;
    if (kc_find_components(b->big, &ob, nullptr, &st) != KC_OK)
      fail(KC_ERR_INVALID, "batch: large graphs: %s", kc_last_error());
    hipLaunchKernelGGL(k_union_stats, dim3(static_cast<unsigned>(ceil_div(b->n_big, kBlock))), dim3(kBlock), 0,
                       b->stream, b->n_big, b->d_big, st.levels, static_cast<uint32_t>(st.sweeps),
                       static_cast<uint32_t>(st.radius_sum), b->dev.glevels, b->dev.gsweeps, b->dev.gradius,
                       b->dev.gstatus, b->dev.gcomp);
    KC_LAUNCH_CHECK();
    if (b->n_union) {
      hipLaunchKernelGGL(k_union_labels, dim3(static_cast<unsigned>(ceil_div(b->n_union, kBlock))), dim3(kBlock), 0,
                         b->stream, b->n_union, b->n_big, b->d_uoff, b->d_big, b->dev.vtx_off,
                         graph_device_labels(b->big), b->dev.labels, b->dev.gcomp);
      KC_LAUNCH_CHECK();
    }
  }
  KC_HIP(hipEventRecord(b->ev1, b->stream));
  KC_HIP(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  KC_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  std::vector<uint32_t> st(b->graphs);
  if (b->graphs) KC_HIP(hipMemcpy(st.data(), b->dev.gstatus, 4ull * b->graphs, hipMemcpyDeviceToHost));
  for (uint32_t g = 0; g < b->graphs; ++g)
    if (st[g]) {
      // a graph that stopped early may have left entries in its hash set
      batch_table_clear(b->table_words, b->table, b->stream);
      KC_HIP(hipStreamSynchronize(b->stream));
      b->ran = false;
      fail(KC_ERR_INVALID, "batch: graph %u stopped early (%s)", g,
           st[g] == 1 ? "level records exceed 10 words per vertex"
                      : st[g] == 2 ? "level cap reached" : "cluster-edge hash set full");
    }
  b->ran = true;
  return ms;
}

}  // namespace

extern "C" {

int kc_batch_create(const kc_batch_host *h, kc_batch **out) {
  return guarded([&] {
    if (!h || !out) fail(KC_ERR_INVALID, "null argument");
    *out = nullptr;
    const uint32_t G = h->graphs;
    if (G && (!h->vtx_off || !h->edge_off || !h->offsets)) fail(KC_ERR_INVALID, "null batch arrays");
    const uint64_t n = G ? h->vtx_off[G] : 0;
    const uint64_t m2 = G ? h->edge_off[G] : 0;
    if (G && (h->vtx_off[0] != 0 || h->edge_off[0] != 0)) fail(KC_ERR_INVALID, "batch: vtx_off[0] and edge_off[0] must be 0");
    if (m2 && !h->targets) fail(KC_ERR_INVALID, "null batch targets");
    if (n >= (1ull << 40)) fail(KC_ERR_INVALID, "batch too large");
    std::vector<uint32_t> noniso(G);
    std::vector<uint64_t> toff(G + 1, 0);
    for (uint32_t g = 0; g < G; ++g) {
      const uint64_t a = h->vtx_off[g], z = h->vtx_off[g + 1];
      if (z < a) fail(KC_ERR_INVALID, "batch: vtx_off not monotone at graph %u", g);
      if (z - a >= 0xFFFFFFFFull) fail(KC_ERR_INVALID, "batch: graph %u has too many vertices", g);
      if (h->offsets[a] != h->edge_off[g] || h->offsets[z] != h->edge_off[g + 1])
        fail(KC_ERR_INVALID, "batch: offsets of graph %u do not match edge_off", g);
      const uint64_t slots = h->edge_off[g + 1] - h->edge_off[g];
      if (h->edge_off[g + 1] < h->edge_off[g] || slots >= (1ull << 31))
        fail(KC_ERR_INVALID, "batch: graph %u needs fewer than 2^31 edge slots", g);
      uint32_t ni = 0;
      for (uint64_t v = a; v < z; ++v) {
        if (h->offsets[v + 1] < h->offsets[v]) fail(KC_ERR_INVALID, "batch: offsets not monotone at vertex %llu", (unsigned long long)v);
        ni += h->offsets[v + 1] > h->offsets[v];
        for (uint64_t e = h->offsets[v]; e < h->offsets[v + 1]; ++e)
          if (h->targets[e] >= z - a)
            fail(KC_ERR_INVALID, "batch: target %u out of range in graph %u", h->targets[e], g);
      }
      noniso[g] = ni;
      uint64_t tw = 64;
      while (tw < 2 * slots) tw <<= 1;
      toff[g + 1] = toff[g] + tw;
    }
    {
      // the contraction reads every cluster edge from both of its ends
      uint32_t bg = 0;
      uint64_t bu = 0, bv = 0;
      if (G && find_one_way_edge_batch(G, h->vtx_off, h->offsets, h->targets, bg, bu, bv))
        fail(KC_ERR_INVALID, "batch: graph %u is not symmetric: edge %llu -> %llu has no reverse", bg,
             (unsigned long long)bu, (unsigned long long)bv);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) fail(KC_ERR_HIP, "no HIP device available");
    std::unique_ptr<kc_batch> b(new kc_batch());
    KC_HIP(hipGetDevice(&b->device));
    hipDeviceProp_t prop;
    KC_HIP(hipGetDeviceProperties(&prop, b->device));
    b->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 1;
    KC_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    KC_HIP(hipEventCreate(&b->ev0));
    KC_HIP(hipEventCreate(&b->ev1));
    KC_HIP(hipHostMalloc(reinterpret_cast<void **>(&b->h_params), sizeof(BatchParams), hipHostMallocDefault));
    b->graphs = G;
    b->n = n;
    b->m2 = m2;
    b->h_vtx_off.assign(h->vtx_off, h->vtx_off + G + 1);
    b->h_edge_off.assign(h->edge_off, h->edge_off + G + 1);
    BatchDev &d = b->dev;
    d.graphs = G;
    uint32_t *order_d = b->dalloc<uint32_t>(G);
    b->order_d = order_d;
    b->d_big = b->dalloc<uint32_t>(G);
    b->d_uoff = b->dalloc<uint64_t>(G + 1);
    b->d_ueoff = b->dalloc<uint64_t>(G + 1);
    uint64_t *vtx = b->dalloc<uint64_t>(G + 1), *eoff = b->dalloc<uint64_t>(G + 1);
    uint64_t *off = b->dalloc<uint64_t>(n + 1);
    uint32_t *col = b->dalloc<uint32_t>(m2);
    uint32_t *ni = b->dalloc<uint32_t>(G);
    uint64_t *tabo = b->dalloc<uint64_t>(G + 1);
    if (G) {
      KC_HIP(hipMemcpy(vtx, h->vtx_off, 8ull * (G + 1), hipMemcpyHostToDevice));
      KC_HIP(hipMemcpy(eoff, h->edge_off, 8ull * (G + 1), hipMemcpyHostToDevice));
      KC_HIP(hipMemcpy(off, h->offsets, 8ull * (n + 1), hipMemcpyHostToDevice));
      if (m2) KC_HIP(hipMemcpy(col, h->targets, 4ull * m2, hipMemcpyHostToDevice));
      KC_HIP(hipMemcpy(ni, noniso.data(), 4ull * G, hipMemcpyHostToDevice));
      KC_HIP(hipMemcpy(tabo, toff.data(), 8ull * (G + 1), hipMemcpyHostToDevice));
    }
    d.order = order_d;
    d.vtx_off = vtx;
    d.edge_off = eoff;
    d.off = off;
    d.col = col;
    d.nonisolated = ni;
    d.table_off = tabo;
    d.labels = b->dalloc<uint32_t>(n);
    d.key = b->dalloc<uint64_t>(n);
    for (int i = 0; i < 2; ++i) {
      d.loff[i] = b->dalloc<uint32_t>(n + G + 1);
      d.lcol[i] = b->dalloc<uint32_t>(m2);
    }
    d.cnt = b->dalloc<uint32_t>(n);
    b->table_words = toff[G];
    b->table = b->dalloc<uint64_t>(b->table_words);
    d.table = b->table;
    d.stack = b->dalloc<uint32_t>(static_cast<uint64_t>(kStackPerVertex) * n);
    d.glevels = b->dalloc<uint32_t>(G);
    d.gsweeps = b->dalloc<uint32_t>(G);
    d.gradius = b->dalloc<uint32_t>(G);
    d.gstatus = b->dalloc<uint32_t>(G);
    d.gcomp = b->dalloc<uint32_t>(G);
    b->queue = b->dalloc<uint32_t>(1);
    d.queue = b->queue;
    b->d_params = b->dalloc<BatchParams>(1);
    d.params = b->d_params;
    batch_table_clear(b->table_words, b->table, b->stream);
    KC_HIP(hipStreamSynchronize(b->stream));
    prepare_split(b.get(), 0);
    *out = b.release();
  });
}

int kc_batch_destroy(kc_batch *b) {
  return guarded([&] { delete b; });
}

int kc_batch_set_tuning(kc_batch *b, const kc_batch_tuning *t) {
  return guarded([&] {
    if (!b || !t) fail(KC_ERR_INVALID, "null argument");
    check_tuning(*t);
    b->tuning = *t;
  });
}

int kc_batch_run(kc_batch *b, const kc_options *opt, uint32_t *labels_out, kc_batch_stats *stats) {
  return guarded([&] {
    if (!b) fail(KC_ERR_INVALID, "null batch");
    kc_options o;
    if (opt)
      o = *opt;
    else
      kc_options_init(&o);
    const auto t0 = std::chrono::steady_clock::now();
    batch_params(b, o);
    const float ms = run_once(b, o);
    if (labels_out && b->n) KC_HIP(hipMemcpy(labels_out, b->dev.labels, 4ull * b->n, hipMemcpyDeviceToHost));
    const auto t1 = std::chrono::steady_clock::now();
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->time_ms = ms;
      stats->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
      stats->n = b->n;
      stats->m = b->m2 / 2;
      stats->graphs = b->graphs;
      stats->tuning = b->tuning;
      std::vector<uint32_t> lv(b->graphs), sw(b->graphs), rd(b->graphs), cc(b->graphs);
      if (b->graphs) {
        KC_HIP(hipMemcpy(lv.data(), b->dev.glevels, 4ull * b->graphs, hipMemcpyDeviceToHost));
        KC_HIP(hipMemcpy(sw.data(), b->dev.gsweeps, 4ull * b->graphs, hipMemcpyDeviceToHost));
        KC_HIP(hipMemcpy(rd.data(), b->dev.gradius, 4ull * b->graphs, hipMemcpyDeviceToHost));
        KC_HIP(hipMemcpy(cc.data(), b->dev.gcomp, 4ull * b->graphs, hipMemcpyDeviceToHost));
      }
      for (uint32_t g = 0; g < b->graphs; ++g) {
        stats->levels_sum += lv[g];
        stats->sweeps_sum += sw[g];
        stats->radius_sum += rd[g];
        stats->components += cc[g];
        stats->max_levels = std::max(stats->max_levels, lv[g]);
      }
    }
  });
}

int kc_batch_autotune(kc_batch *b, const kc_options *opt, kc_batch_tuning *chosen) {
  return guarded([&] {
    if (!b) fail(KC_ERR_INVALID, "null batch");
    kc_options o;
    if (opt)
      o = *opt;
    else
      kc_options_init(&o);
    const kc_batch_tuning keep = b->tuning;
    auto time_of = [&](const kc_batch_tuning &t) {
      b->tuning = t;
      batch_params(b, o);
      run_once(b, o);  // warm
      const float ms = run_once(b, o);
      return std::min(ms, run_once(b, o));
    };
    // the split size first, at the current launch shape, then the shape
    kc_batch_tuning best = keep;
    best.use_graph = 0;
    float best_ms = 0.f;
    bool have = false;
    const uint32_t splits[] = {0u, 1u << 12, 1u << 14, 1u << 16};
    for (uint32_t sp : splits) {
      kc_batch_tuning t = best;
      t.split_vertices = sp;
      const float ms = time_of(t);
      if (!have || ms < best_ms) {
        best_ms = ms;
        best = t;
        have = true;
      }
    }
    const int blocks[] = {256, 512, 1024};
    const int gpcs[] = {1, 2, 4, 8};
    for (int blk : blocks)
      for (int gpc : gpcs) {
        if (blk * gpc > 2048) continue;  // at most 32 waves per compute unit
        kc_batch_tuning t{blk, gpc, 0, best.split_vertices};
        const float ms = time_of(t);
        if (ms < best_ms) {
          best_ms = ms;
          best = t;
        }
      }
    best.use_graph = keep.use_graph;
    b->tuning = best;
    if (chosen) *chosen = best;
  });
}

int kc_batch_graph_stats(const kc_batch *b, uint32_t *levels, uint32_t *sweeps, uint32_t *radius) {
  return guarded([&] {
    if (!b) fail(KC_ERR_INVALID, "null batch");
    if (!b->ran) fail(KC_ERR_INVALID, "batch: no completed run");
    if (!b->graphs) return;
    if (levels) KC_HIP(hipMemcpy(levels, b->dev.glevels, 4ull * b->graphs, hipMemcpyDeviceToHost));
    if (sweeps) KC_HIP(hipMemcpy(sweeps, b->dev.gsweeps, 4ull * b->graphs, hipMemcpyDeviceToHost));
    if (radius) KC_HIP(hipMemcpy(radius, b->dev.gradius, 4ull * b->graphs, hipMemcpyDeviceToHost));
  });
}

int kc_batch_find_components(const kc_batch_host *h, const kc_options *opt, uint32_t *labels_out,
                             kc_batch_stats *stats) {
  kc_batch *b = nullptr;
  int rc = kc_batch_create(h, &b);
  if (rc != KC_OK) return rc;
  rc = kc_batch_run(b, opt, labels_out, stats);
  const std::string err = rc == KC_OK ? std::string() : std::string(kc_last_error());
  kc_batch_destroy(b);
  if (rc != KC_OK) set_error("%s", err.c_str());
  return rc;
}

}  // extern "C"
