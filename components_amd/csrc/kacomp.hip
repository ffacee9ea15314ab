// libkacomp C ABI: graph handle, level loop of exponential contraction and its
// variants, undo and canonical labels.  One HIP stream per graph handle; every
// call is synchronous on return.
//
// Level loop (one GPU, the P -> n limit of the reference where every vertex is
// an interface vertex, so the contraction runs on the whole graph):
//   RunContractionLP           lib/components/exponential_contraction.h:439-628
//   ExponentialContraction     lib/contraction/dynamic_contraction.h:62-136
//   UndoContraction            lib/contraction/dynamic_contraction.h:568-658
//   ApplyToLocalComponents     lib/components/exponential_contraction.h:876-889
// Ghost exchange (lib/communication) has no counterpart: neighbours read the
// key array directly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "csr_build.h"
#include "deviates.h"
#include "kc_common.h"
#include "kernels.h"
#include "learned.h"

namespace kc {

// ---------------------------------------------------------------------------
// Errors.
static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

void fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  throw Failure{code};
}

// ---------------------------------------------------------------------------
// Arena.
Arena::~Arena() {
  for (auto &c : chunks_) (void)hipFree(c.first);
}

void *Arena::alloc(size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  while (cur_ < chunks_.size()) {
    if (used_ + bytes <= chunks_[cur_].second) {
      void *p = chunks_[cur_].first + used_;
      used_ += bytes;
      return p;
    }
    ++cur_;
    used_ = 0;
  }
  size_t sz = std::max<size_t>(bytes, size_t(64) << 20);
  char *p = nullptr;
  if (hipMalloc(&p, sz) != hipSuccess) fail(KC_ERR_OOM, "device arena: cannot allocate %zu bytes", sz);
  chunks_.emplace_back(p, sz);
  cur_ = chunks_.size() - 1;
  used_ = bytes;
  return p;
}

size_t Arena::reserved() const {
  size_t s = 0;
  for (auto &c : chunks_) s += c.second;
  return s;
}

// ---------------------------------------------------------------------------
// Kernel timer.
KernelTimer::~KernelTimer() {
  for (auto e : pool_) (void)hipEventDestroy(e);
}

void KernelTimer::begin_run(bool enabled, hipStream_t s, int only_cls) {
  enabled_ = enabled;
  only_ = only_cls;
  stream_ = s;
  recs_.clear();
  next_ev_ = 0;
}

int KernelTimer::start(int cls) {
  int ev = -1;
  if (enabled_ && (only_ < 0 || cls == only_)) {
    while (next_ev_ + 2 > pool_.size()) {
      hipEvent_t e;
      KC_HIP(hipEventCreate(&e));
      pool_.push_back(e);
    }
    ev = static_cast<int>(next_ev_);
    next_ev_ += 2;
    KC_HIP(hipEventRecord(pool_[ev], stream_));
  }
  recs_.push_back({cls, ev, 0});
  return static_cast<int>(recs_.size() - 1);
}

void KernelTimer::stop(int slot, uint64_t alg_bytes) {
  Rec &r = recs_[slot];
  r.bytes += alg_bytes;
  if (r.ev >= 0) KC_HIP(hipEventRecord(pool_[r.ev + 1], stream_));
}

void KernelTimer::retag(int slot, int cls, uint64_t alg_bytes) {
  recs_[slot].cls = cls;
  recs_[slot].bytes = alg_bytes;
}

void KernelTimer::finish(kc_stats *st) {
  if (!st) return;
  bool seen[KC_KERNEL_CLASSES];
  for (int i = 0; i < KC_KERNEL_CLASSES; ++i) {
    st->kernel_ms[i] = 0.0;
    st->kernel_launches[i] = 0;
    st->kernel_bytes[i] = 0;
    st->kernel_first_ms[i] = 0.0;
    st->kernel_first_bytes[i] = 0;
    seen[i] = false;
  }
  if (enabled_) KC_HIP(hipStreamSynchronize(stream_));
  for (auto &r : recs_) {  // launches without events add no time
    st->kernel_launches[r.cls]++;
    st->kernel_bytes[r.cls] += r.bytes;
    if (r.ev >= 0) {
      float ms = 0.f;
      KC_HIP(hipEventElapsedTime(&ms, pool_[r.ev], pool_[r.ev + 1]));
      st->kernel_ms[r.cls] += ms;
      if (!seen[r.cls]) {
        seen[r.cls] = true;
        st->kernel_first_ms[r.cls] = ms;
        st->kernel_first_bytes[r.cls] = r.bytes;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Deviate threshold table (host, glibc log as in the reference).
static uint64_t truncated_deviate(double w, double beta) {
  return static_cast<uint64_t>(-std::log(w) / beta);
}

std::vector<double> build_deviate_table(double beta) {
  const double wmin = 0x1p-53;  // smallest 1-u the canonical draw can produce
  const uint64_t kmax = truncated_deviate(wmin, beta);
  std::vector<double> t(kmax);
  uint64_t lo_bits, hi_bits;
  for (uint64_t k = 1; k <= kmax; ++k) {
    double lo = wmin, hi = 1.0;  // F(lo) >= k, F(hi) = 0 < k
    std::memcpy(&lo_bits, &lo, 8);
    std::memcpy(&hi_bits, &hi, 8);
    while (hi_bits - lo_bits > 1) {
      uint64_t mid = lo_bits + (hi_bits - lo_bits) / 2;
      double w;
      std::memcpy(&w, &mid, 8);
      if (truncated_deviate(w, beta) >= k)
        lo_bits = mid;
      else
        hi_bits = mid;
    }
    std::memcpy(&t[k - 1], &lo_bits, 8);
  }
  return t;
}

}  // namespace kc

using namespace kc;

// ---------------------------------------------------------------------------
// Graph handle.
struct LevelRec {
  bool stall = false;  // no key changed: the contraction is the identity and is skipped
  const uint32_t *idx = nullptr;  // level stored in another order: position -> level index
  uint64_t n = 0, m2 = 0;
  uint32_t *lab = nullptr, *newidx = nullptr, *gid = nullptr;
  uint64_t sweeps = 0;
  uint32_t radius = 0;
  uint64_t *keys_copy = nullptr;  // trace only
};

struct kc_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t n = 0, m2 = 0, non_isolated = 0;
  // input CSR
  uint64_t *off0 = nullptr;
  uint32_t *col0 = nullptr;
  // level 0 of the contraction: the n0 vertices of degree > 0 in id order,
  // gid0 their input ids (null when no vertex is isolated: then off0c/col0c
  // are off0/col0)
  uint64_t n0 = 0;
  uint32_t *gid0 = nullptr;
  uint64_t *off0c = nullptr;
  uint32_t *col0c = nullptr;
  // the same level 0 stored in degree-descending order (hub keys stay in
  // cache): idx0 position -> level index, pos0 its inverse, gidp0 position ->
  // input id, off0p/col0p the CSR with positions (null when not built)
  uint32_t *idx0 = nullptr, *pos0 = nullptr, *gidp0 = nullptr;
  uint64_t *off0p = nullptr;
  uint32_t *col0p = nullptr;
  // work buffers sized for level 0
  uint64_t *key = nullptr, *key2 = nullptr;
  uint32_t *deg_inter = nullptr, *lists = nullptr, *hub_inter = nullptr;
  HubChunk *hubs = nullptr;
  uint32_t *flists = nullptr;  // bins of the unsettled vertices (time-stepped schedule)
  HubChunk *fhubs = nullptr;
  // level 0's bins depend on its degrees alone: built by the first call, kept
  uint32_t *l0_lists = nullptr;
  HubChunk *l0_hubs = nullptr;
  Bins l0_bins;
  // the same bins with every list in level-index order, for round 0 of the
  // degree-ordered level 0 (built when first used, dropped with l0_bins)
  uint32_t *l0i_lists = nullptr;
  HubChunk *l0i_hubs = nullptr;
  Bins l0i_bins;
  bool l0i_valid = false;
  bool l0i_hubs_first = false;  // the hub chunks' rows carry small indices: their section goes first
  BinCfg l0_cfg{0, 0, 0};
  const uint64_t *l0_off = nullptr;
  uint64_t hubs_cap = 0;
  uint32_t *slots = nullptr, *seg_of_cid = nullptr, *uniq = nullptr, *pos_of = nullptr, *hub_pos = nullptr;
  uint64_t *hub_prefix = nullptr;  // exclusive prefix of the hub chunks' counts
  uint64_t *row_cstart = nullptr;  // chunk plan of the listed buckets' rows (column gather)
  uint32_t *crank = nullptr, *cid = nullptr;  // dense cluster ids: rank by label, id per vertex
  uint32_t *sk = nullptr, *sv = nullptr, *sk2 = nullptr, *sv2 = nullptr;  // (label, vertex) sort buffers
  uint64_t *seg_off = nullptr, *sort_scratch = nullptr;
  uint32_t *rep_a = nullptr, *rep_b = nullptr, *minid = nullptr, *labels = nullptr;
  uint64_t *scan_scratch = nullptr;
  uint64_t *off_pp[2] = {nullptr, nullptr};
  uint32_t *col_pp[2] = {nullptr, nullptr};
  uint32_t *dl_medium = nullptr, *dl_huge = nullptr, *hcount = nullptr;
  uint32_t *huge_tables = nullptr;
  uint64_t huge_cap = 0;
  uint32_t *huge_bits = nullptr;  // bitmaps of the huge buckets over the label space
  uint64_t huge_bits_cap = 0;
  uint8_t *huge_maps = nullptr;  // byte maps of the huge buckets over the label space
  uint64_t huge_maps_cap = 0;
  HugeChunk *d_chunks = nullptr;
  uint64_t chunks_cap = 0;
  uint64_t *d_huge_info = nullptr;
  uint64_t huge_info_cap = 0;
  int *flags = nullptr;
  uint32_t *small = nullptr;  // [0..3] bin counters, [4] radius, [5] roots, [6] hubs found, [7] packed keys do not fit
  uint64_t *sizes = nullptr;  // [0] n_next [1] P [2] m2_next [3..5] dedup counters [8] N [9] clusters
  // pinned host mirrors
  uint64_t *h_sizes = nullptr;
  uint32_t *h_small = nullptr;
  int *h_flags = nullptr;
  // deviate table
  double *d_table = nullptr;
  uint32_t table_k = 0;
  double table_beta = -1.0;
  uint64_t table_cap = 0;
  // run state
  Arena arena;
  KernelTimer timer;
  std::vector<LevelRec> levels;
  bool base_case_used = false;
  uint64_t base_case_rounds = 0;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  uint32_t components = 0;
  bool have_labels = false;
  // learned centre proposals: the model, and per-vertex buffers sized for level 0
  kc_mlp *mlp = nullptr;
  kc_mlp *collect = nullptr;  // set while kc_mlp_collect records a plain run
  uint64_t collect_stride = 1;
  float *feat = nullptr, *target = nullptr;
  uint32_t *ldeg = nullptr, *csize = nullptr;
  uint8_t *delay = nullptr;
  // Scores of the input's level 0: its features (round index 0) and so its
  // scores depend on the graph and the model alone, not on the seed or alpha;
  // kept for the model and storage layout they were computed for.
  float *l0_score = nullptr;
  const kc_mlp *l0_score_mlp = nullptr;
  uint64_t l0_score_version = 0;
  const uint64_t *l0_score_off = nullptr;
  uint64_t learned_cap = 0;  // vertices the learned buffers hold
  BinCfg tuned{0, 0, 0};     // thresholds kept by kc_graph_autotune (tmax 0: none)
  // the degree-ordered level 0's first kHotClusters positions hold at least
  // kHotShareFactor times their proportional share of its edge slots
  bool l0_hot_skewed = false;
  // Pieces of the degree-ordered level 0's rows above the 16-lane bin for the
  // owner-split relabel (one per owner group of a row, cut every wmax slots),
  // their counts, offsets in their row's run and prefix, the piece indices of
  // each owner part (part p: l0_part_list[l0_part_at[p] .. l0_part_at[p+1]));
  // built by the first call that needs them, for the bins they were cut with.
  HubChunk *l0_pieces = nullptr;
  uint32_t *l0_part_list = nullptr, *l0_piece_inter = nullptr, *l0_piece_pos = nullptr;
  uint64_t *l0_piece_prefix = nullptr;
  uint64_t l0_npieces = 0;
  uint32_t l0_part_at[kParts + 2] = {};  // [kParts .. kParts+1): the shared pieces
  BinCfg l0_piece_cfg{0, 0, 0};
  uint64_t l0_piece_split = 0;  // rows above this many slots cut by owner part
  bool l0_pieces_valid = false;
  // level-sized work buffers hold levels of up to cap_n vertices and cap_m2
  // edge slots (the input's level 0, or a larger replicated one)
  uint64_t cap_n = 0, cap_m2 = 0;
  std::vector<void *> owned;       // freed with the handle
  std::vector<void *> work_owned;  // the level-sized work set, freed when it grows

  template <class T>
  T *alloc_into(std::vector<void *> &list, uint64_t count) {
    void *p = nullptr;
    size_t bytes = std::max<uint64_t>(count, 1) * sizeof(T) + 64;
    if (hipMalloc(&p, bytes) != hipSuccess) fail(KC_ERR_OOM, "hipMalloc of %zu bytes failed", bytes);
    list.push_back(p);
    return static_cast<T *>(p);
  }
  template <class T>
  T *dalloc(uint64_t count) {
    return alloc_into<T>(owned, count);
  }
  template <class T>
  T *walloc(uint64_t count) {
    return alloc_into<T>(work_owned, count);
  }
  // Replaces a grown buffer: the stream is drained, the old one freed.
  template <class T>
  void drealloc(T *&p, uint64_t count) {
    if (p) {
      (void)hipStreamSynchronize(stream);
      auto it = std::find(owned.begin(), owned.end(), static_cast<void *>(p));
      if (it != owned.end()) owned.erase(it);
      (void)hipFree(p);
      p = nullptr;
    }
    p = dalloc<T>(count);
  }
  void free_work() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void *p : work_owned) (void)hipFree(p);
    work_owned.clear();
  }
  ~kc_graph() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void *p : work_owned) (void)hipFree(p);
    for (void *p : owned) (void)hipFree(p);
    if (h_sizes) (void)hipHostFree(h_sizes);
    if (h_small) (void)hipHostFree(h_small);
    if (h_flags) (void)hipHostFree(h_flags);
    if (ev_begin) (void)hipEventDestroy(ev_begin);
    if (ev_end) (void)hipEventDestroy(ev_end);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

constexpr uint32_t kMaxFlags = 1u << 16;
constexpr uint64_t kOrderMinVertices = 1u << 16;  // smaller level-0 graphs keep the input order
// Owner-split relabel: the degree-ordered level 0's rows of more slots than
// this are cut by the owner part of their targets (shorter rows above the
// 16-lane bin stay whole)
constexpr uint64_t kSplitRowsAbove = 4096;
// Hot-first cluster ids (and the owner-split relabel, which needs them) when the
// first kHotClusters positions of the degree-ordered level 0 hold at least
// this many times their proportional share of the edge slots (R-MAT 24: 70
// times, the hyperbolic graph of 2^24 vertices: 23; G(n,m) 2^20, the 3D grid
// 256^3 and the 2D geometric graph of 2^24 vertices: 1.8 to 2.6)
constexpr double kHotShareFactor = 4.0;
constexpr uint64_t kLevelHotRows = 28672;  // labels the one-pass relabel serves from LDS (32-bit)
// -bfs: pull sweeps finish a level once at most n/8 rows are pending (3D grid 256^3: 7.01 -> 6.90 ms
// against n/32; n/4 took 7.22, R-MAT 24 and the others alike)
constexpr int kTailShift = 3;
constexpr uint64_t kPackedSweepsMin = 1u << 16;  // levels at least this large sweep packed keys
constexpr uint64_t kPullRoundsMin = 1u << 16;  // -bfs: levels at least this large choose round directions
constexpr uint32_t kDirectionRounds = 6;        // rounds 1 .. this may be pulled
constexpr uint64_t kDirectionMinDegree = 8;     // ... on levels of this average degree or more (sparse,
                                                // high-diameter levels keep batched pushes)
constexpr uint64_t kHugeBitmapBytes = 2ull << 30;  // above this the huge buckets use hash tables
constexpr uint64_t kHugeMapBytes = 8ull << 30;     // byte maps of the largest huge buckets
constexpr uint64_t kMapMinShare = 32;              // byte maps from C / 32 elements on
constexpr uint64_t kHubPullMin = 65536;            // smallest degree of a hub that pulls round 0


void sync(kc_graph *g) { KC_HIP(hipStreamSynchronize(g->stream)); }

void d2h(kc_graph *g, void *dst, const void *src, size_t bytes) {
  KC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g->stream));
}

void ensure_table(kc_graph *g, double beta) {
  if (g->table_beta == beta && g->d_table) return;
  if (!(beta > 1e-6) || !(beta < 1e6)) fail(KC_ERR_INVALID, "beta must be in (1e-6, 1e6), got %g", beta);
  std::vector<double> t = build_deviate_table(beta);
  if (t.size() + 1 > g->table_cap) {
    g->table_cap = t.size() + 1;
    g->drealloc(g->d_table, g->table_cap);
  }
  if (!t.empty())
    KC_HIP(hipMemcpy(g->d_table, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
  g->table_k = static_cast<uint32_t>(t.size());
  g->table_beta = beta;
}

// The options' thresholds, or the handle's tuned ones when the options leave
// all three at 0.
BinCfg bin_cfg(const kc_graph *g, const kc_options *o) {
  if (g && g->tuned.tmax && o->sweep_group_small <= 0 && o->sweep_group_mid <= 0 && o->sweep_group_large <= 0)
    return g->tuned;
  return bin_config(o->sweep_group_small, o->sweep_group_mid, o->sweep_group_large);
}

// Bins of the level CSR (one sync to size the binned launches).  With key,
// only the vertices at key distance >= tmin, into the second set of lists.
// Level 0's full bins go to their own lists and are reused by later calls
// with the same thresholds (no launch, no sync).
Bins make_bins(kc_graph *g, uint64_t n, const uint64_t *off, const BinCfg &cfg, const uint64_t *key = nullptr,
               uint32_t tmin = 0) {
  const bool level0 = !key && n == g->n0 && (off == g->off0p || off == g->off0c);
  if (level0 && g->l0_off == off && g->l0_cfg.tmax == cfg.tmax && g->l0_cfg.smax == cfg.smax &&
      g->l0_cfg.wmax == cfg.wmax)
    return g->l0_bins;
  if (level0 && !g->l0_lists) {
    g->l0_lists = g->walloc<uint32_t>(3 * g->cap_n);
    g->l0_hubs = g->walloc<HubChunk>(g->hubs_cap);
  }
  if (level0) g->l0i_valid = false;
  // three sections of cap_n entries each: a level holds at most cap_n vertices
  // (a replicated level 0 can hold more than the input's n)
  if (n > g->cap_n) fail(KC_ERR_INVALID, "internal: level of %llu vertices above the work set's %llu",
                         static_cast<unsigned long long>(n), static_cast<unsigned long long>(g->cap_n));
  Bins b;
  uint32_t *lists = key ? g->flists : level0 ? g->l0_lists : g->lists;
  b.listT = lists;
  b.listS = lists + g->cap_n;
  b.listW = lists + 2 * g->cap_n;
  b.hubs = key ? g->fhubs : level0 ? g->l0_hubs : g->hubs;
  {
    TimedScope ts(g->timer, KC_K_BIN, (key ? 24 : 16) * n + 8);
    bin_vertices(n, off, cfg, b.listT, b.listS, b.listW, b.hubs, g->small, key ? nullptr : g->deg_inter,
                 g->scan_scratch, g->stream, key, tmin);
  }
  d2h(g, g->h_small, g->small, 4 * sizeof(uint32_t));
  sync(g);
  b.nT = g->h_small[0];
  b.nS = g->h_small[1];
  b.nW = g->h_small[2];
  b.nH = g->h_small[3];
  if (b.nH > g->hubs_cap) fail(KC_ERR_INVALID, "hub chunk overflow (%u > %llu)", b.nH, (unsigned long long)g->hubs_cap);
  finish_bins(b);
  if (level0) {
    g->l0_bins = b;
    g->l0_cfg = cfg;
    g->l0_off = off;
  }
  return b;
}

uint64_t sweep_bytes(uint64_t n, uint64_t m2, bool count) {
  // offsets, neighbour ids, neighbour keys, own key read + write, inter count
  return 8 * (n + 1) + 12 * m2 + 16 * n + (count ? 4 * n : 0);
}

// Runs in-place sweeps (and optionally pointer jumps) until no key changes.
// Sweeps are launched in speculative batches; a sweep whose predecessor
// changed nothing returns at once.  Returns the number of sweeps that did work
// (the last of them found no change).
// packed (non-null: plain in-place sweeps without counts or jumps): the
// sweeps run over these packed keys (pack_keys) instead of key.
uint64_t sweep_to_fixpoint(kc_graph *g, uint64_t n, uint64_t m2, const uint64_t *off, const uint32_t *col,
                           uint64_t *key, uint64_t hop, const Bins &b, bool count, bool with_jump,
                           uint32_t &flag_base, bool book_bytes = true, uint64_t block = 0,
                           uint32_t *packed = nullptr) {
  uint64_t done = 0;
  const uint32_t batch = n < (1u << 16) ? 8u : 3u;
  // (packed: 4-byte keys for the neighbours and the row's own)
  const uint64_t bytes = !book_bytes ? 0 : packed ? 8 * (n + 1) + 8 * m2 + 8 * n : sweep_bytes(n, m2, count);
  while (true) {
    if (flag_base + batch + 2 >= kMaxFlags) {
      fill_u32(reinterpret_cast<uint32_t *>(g->flags), kMaxFlags, 0, g->stream);
      flag_base = 0;
    }
    int slots[8];
    for (uint32_t i = 0; i < batch; ++i) {
      const uint32_t f = flag_base + i;
      // the first sweep of a batch always runs: the previous batch ended with a change
      const int *prev = i == 0 ? nullptr : &g->flags[f - 1];
      slots[i] = g->timer.start(KC_K_SWEEP);
      // hub chunks first, then wave, 16-lane and one-lane rows: the many
      // short rows then read the keys the hubs lowered in the same sweep
      // (R-MAT 24, sweeps schedule: 11.8 -> 11.0 sweeps, 24.7 -> 21.9 ms)
      if (packed)
        sweep_packed(off, col, packed, b, &g->flags[f], prev, g->stream, true);
      else
        sweep(block ? kSweepBlock : kSweepInPlace, count, off, col, key, key, hop, b, g->deg_inter, g->hub_inter,
              &g->flags[f], prev, g->stream, block, true);
      g->timer.stop(slots[i], bytes);
      if (with_jump) {
        TimedScope js(g->timer, KC_K_JUMP, 16 * n);
        jump(n, key, &g->flags[f], g->stream);
      }
    }
    d2h(g, g->h_flags, &g->flags[flag_base], batch * sizeof(int));
    sync(g);
    uint32_t i = 0;
    while (i < batch && g->h_flags[i] != 0) ++i;
    const uint32_t worked = i < batch ? i + 1 : batch;
    // launches after the last change returned at once: they keep their class
    // (so launch counts match a kernel trace) but carry no bytes
    for (uint32_t j = worked; j < batch; ++j) g->timer.retag(slots[j], KC_K_SWEEP, 0);
    done += worked;
    flag_base += batch;
    if (i < batch) break;
  }
  return done;
}

// Dedup of buckets [seg_off[a2], seg_off[a2+1]) for a2 < nb: small buckets
// in wave tables, medium ones in block tables, huge ones against byte maps or
// bitmaps over the label space (mode KC_DEDUP_AUTO: byte maps while they fit
// their cap, then bitmaps) or in global tables processed in fixed-size chunks
// planned on the host.  The medium and huge lists (n_medium, n_huge) come
// from classify_buckets; total = elements of all buckets.  Returns the timer
// slot of the first dedup launch (its algorithmic bytes are booked by the
// caller).  `scratch` serves the scans (the handle's own scan scratch does not
// exist yet while a generated graph's CSR is built).
int dedup_buckets(kc_graph *g, uint64_t nb, const uint64_t *seg_off, const uint32_t *bucket, uint32_t *out,
                  uint32_t *uniq, const uint32_t *medium, const uint32_t *huge, uint64_t n_medium, uint64_t n_huge,
                  uint64_t total, uint64_t label_space, uint64_t *scratch, int mode = KC_DEDUP_AUTO) {
  hipStream_t s = g->stream;
  int first;
  {
    TimedScope ts(g->timer, KC_K_DEDUP, 0);
    first = ts.slot;
    dedup_small(nb, seg_off, bucket, out, uniq, s);
  }
  if (n_medium) {
    TimedScope ts(g->timer, KC_K_DEDUP, 0);
    dedup_medium(static_cast<uint32_t>(n_medium), medium, seg_off, bucket, out, uniq, s);
  }
  if (!n_huge) return first;
  if (3 * n_huge + 6 > g->huge_info_cap) {
    g->huge_info_cap = 4 * n_huge + 64;
    g->drealloc(g->d_huge_info, g->huge_info_cap);
  }
  // Bitmaps over the label space when they fit their cap, the buckets of at
  // least map_min elements against byte maps instead: those cost two passes
  // over C bytes (the emit reads the map and clears it) but no atomics, the
  // bitmaps one device atomic per distinct value (about 0.12 ns each against
  // 0.4 ps per map byte), so the maps pay from about C/128 distinct values
  // on; a bucket's distinct values are fewer than its elements, and on R-MAT
  // 24 the step is fastest with the maps from C/32 elements on (flat between
  // C/16 and C/64).  Their number is at most total / map_min, which bounds the
  // maps' memory before any size is read back; map_min rises until that bound
  // fits the maps' cap.
  const uint64_t words = (label_space + 31) / 32;
  const bool bits_fit = n_huge * words * 4 <= kHugeBitmapBytes;
  const uint64_t stride = huge_map_stride(label_space);
  uint64_t map_min = ~0ull, n_maps = 0;
  if ((mode == KC_DEDUP_AUTO || mode == KC_DEDUP_BYTEMAPS || mode == KC_DEDUP_MAPS_TABLES) && label_space) {
    map_min = mode == KC_DEDUP_BYTEMAPS ? kDedupBlockCap + 1
                                        : std::max<uint64_t>(kDedupBlockCap + 1, label_space / kMapMinShare);
    const uint64_t cap_maps = kHugeMapBytes / stride;
    while (map_min != ~0ull && std::min(n_huge, total / map_min) > cap_maps)
      map_min = cap_maps ? total / cap_maps + 1 : ~0ull;
    if (map_min != ~0ull) n_maps = std::min(n_huge, total / map_min);
    if (!n_maps) map_min = ~0ull;
  }
  if (n_maps * stride > g->huge_maps_cap) {  // maps start zero and every call leaves them zero
    g->huge_maps_cap = std::min(kHugeMapBytes, n_maps * stride + (n_maps * stride >> 2));
    g->drealloc(g->huge_maps, g->huge_maps_cap);
    fill_u32(reinterpret_cast<uint32_t *>(g->huge_maps), g->huge_maps_cap / 4, 0u, s);
  }
  if (mode == KC_DEDUP_BYTEMAPS || mode == KC_DEDUP_BITMAPS || (mode == KC_DEDUP_AUTO && bits_fit)) {
    if (map_min > kDedupBlockCap + 1 && n_huge * words > g->huge_bits_cap) {
      g->huge_bits_cap = n_huge * words + (n_huge * words >> 2);
      g->drealloc(g->huge_bits, g->huge_bits_cap);
    }
    TimedScope ts(g->timer, KC_K_DEDUP, 0);
    dedup_huge_mixed(static_cast<uint32_t>(n_huge), huge, seg_off, g->m2 / kBitsChunk + n_huge, map_min, n_maps, true,
                     true, g->d_huge_info, g->huge_bits, words, g->huge_maps, label_space, bucket, out, uniq, scratch,
                     s);
    return first;
  }
  // Global hash tables (KC_DEDUP_TABLES, or the bitmaps over their cap): in
  // AUTO (and MAPS_TABLES) mode the buckets of at least map_min elements still
  // go to byte maps after the tables, the others to tables sized by their
  // bucket (a bitmap costs C/8 bytes however small the bucket).
  if (mode == KC_DEDUP_TABLES) map_min = ~0ull, n_maps = 0;
  huge_sizes(static_cast<uint32_t>(n_huge), huge, seg_off, g->d_huge_info, s);
  std::vector<uint64_t> info(2 * n_huge);
  std::vector<uint32_t> a2s(n_huge);
  KC_HIP(hipMemcpyAsync(info.data(), g->d_huge_info, 16 * n_huge, hipMemcpyDeviceToHost, s));
  KC_HIP(hipMemcpyAsync(a2s.data(), huge, 4 * n_huge, hipMemcpyDeviceToHost, s));
  sync(g);
  std::vector<HugeChunk> elem, slot;
  uint64_t huge_slots = 0;  // tables of at least twice their bucket, one after another
  for (uint64_t h = 0; h < n_huge; ++h) {
    const uint64_t size = info[2 * h + 1];
    if (size >= map_min) continue;  // a byte map's bucket: its count stays 0 until the maps add to it
    uint64_t tsz = 1;
    while (tsz < 2 * size) tsz <<= 1;
    HugeChunk c{info[2 * h], huge_slots, 0, 0, static_cast<uint32_t>(tsz - 1), static_cast<uint32_t>(h), a2s[h], 0};
    huge_slots += tsz;
    for (uint64_t lo = 0; lo < size; lo += kHugeChunk) {
      c.lo = lo;
      c.hi = std::min(size, lo + kHugeChunk);
      elem.push_back(c);
    }
    for (uint64_t lo = 0; lo < tsz; lo += kHugeChunk) {
      c.lo = lo;
      c.hi = std::min(tsz, lo + kHugeChunk);
      slot.push_back(c);
    }
  }
  if (elem.size() + slot.size() > g->chunks_cap) {
    g->chunks_cap = elem.size() + slot.size() + 1024;
    g->drealloc(g->d_chunks, g->chunks_cap);
  }
  KC_HIP(hipMemcpyAsync(g->d_chunks, elem.data(), elem.size() * sizeof(HugeChunk), hipMemcpyHostToDevice, s));
  KC_HIP(hipMemcpyAsync(g->d_chunks + elem.size(), slot.data(), slot.size() * sizeof(HugeChunk),
                        hipMemcpyHostToDevice, s));
  if (huge_slots > g->huge_cap) {
    g->huge_cap = huge_slots + (huge_slots >> 2);
    g->drealloc(g->huge_tables, g->huge_cap);
  }
  if (n_huge > g->huge_info_cap) fail(KC_ERR_INVALID, "huge bucket list overflow");
  fill_u32(g->huge_tables, huge_slots, kNone, s);
  uint32_t *hcount = reinterpret_cast<uint32_t *>(g->d_huge_info);  // sizes already copied out
  fill_u32(hcount, n_huge, 0, s);
  TimedScope ts(g->timer, KC_K_DEDUP, 0);
  dedup_huge(g->d_chunks, static_cast<uint32_t>(elem.size()), g->d_chunks + elem.size(),
             static_cast<uint32_t>(slot.size()), static_cast<uint32_t>(n_huge), huge, g->huge_tables, bucket, out, uniq,
             hcount, s);
  if (n_maps)  // the rest against byte maps (uniq of their buckets left at 0 by the tables' counts, if any ran)
    dedup_huge_mixed(static_cast<uint32_t>(n_huge), huge, seg_off, g->m2 / kBitsChunk + n_huge, map_min, n_maps, false,
                     elem.empty(), g->d_huge_info, g->huge_bits, words, g->huge_maps, label_space, bucket, out, uniq, scratch,
                     s);
  // the chunk vectors die here: wait until the uploads have been consumed
  sync(g);
  return first;
}

// Level 0's cached bins in level-index order (see l0i_bins).
Bins round0_bins(kc_graph *g, const Bins &b);

// Time-stepped schedule until no vertex lies above the current round, then
// one counting pass for the cluster-edge counts.  Returns the rounds that did
// work (the reported "sweeps" of this schedule).
uint64_t rounds_to_fixpoint(kc_graph *g, uint64_t n, uint64_t m2, const uint64_t *off, const uint32_t *col,
                            uint64_t *key, const Bins &level_bins, const BinCfg &cfg, uint32_t &flag_base,
                            int tail_shift, bool choose_direction, bool index_round0) {
  uint64_t done = 0;
  uint32_t t = 0;
  bool deciding = choose_direction;
  Bins b = level_bins;
  uint32_t next_rebin = 4;
  int first_slot = -1;
  while (true) {
    if (deciding && t >= 1) {
      // Direction of round t (as in direction-optimising BFS): pushed by the
      // rows at distance t, or pulled by the rows above it when those hold
      // fewer edge slots (as in round 1 of R-MAT, where most rows sit at
      // distance 1 after round 0).  A pull lets every vertex above t take the
      // best neighbour key + 1 in place: the keys up to distance t are final,
      // so every vertex whose final distance is t + 1 ends with its final key
      // -- what round t guarantees -- and the others keep upper bounds that
      // later rounds lower.
      // (the edge slots at and above t were counted after round t - 1 and
      // came back with its flag)
      if (g->h_sizes[13] < g->h_sizes[12]) {
        b = make_bins(g, n, off, cfg, key, t + 1);
        if (flag_base + 3 >= kMaxFlags) {
          fill_u32(reinterpret_cast<uint32_t *>(g->flags), kMaxFlags, 0, g->stream);
          flag_base = 0;
        }
        {
          TimedScope ts(g->timer, KC_K_SWEEP, 0);
          sweep(kSweepInPlace, false, off, col, key, key, kOneHop, b, g->deg_inter, g->hub_inter,
                &g->flags[flag_base], nullptr, g->stream);
        }
        ++flag_base;
        ++done;
        ++t;
        deciding = false;  // the rows left above t are few; the rounds go on as pushes
        while (next_rebin <= t) next_rebin *= 2;
      } else if (t >= kDirectionRounds) {
        deciding = false;
      }
    }
    // the first rounds carry most of the pushes; later ones (over the
    // re-binned unsettled vertices) are cheap, so they go in longer
    // speculative batches with one synchronisation each
    // (one round at a time while the direction is chosen: the choice reads
    // the previous round's result)
    const uint32_t batch = deciding ? 1u : t < 8 ? 4u : 16u;
    if (t >= next_rebin) {
      // drop the settled vertices from the rounds' bins (distances below t)
      b = make_bins(g, n, off, cfg, key, t);
      next_rebin *= 2;
      const uint64_t pending = uint64_t(b.nT) + b.nS + b.nW + b.nH;
      if (tail_shift >= 0 && (pending << tail_shift) <= n) {
        // Few vertices left: finish with in-place pull sweeps over them.  The
        // settled vertices have pushed their final keys already, so the keys
        // still to change follow from the pending rows alone, and the sweeps
        // reach the same fixed point in as many sweeps as the longest chain
        // of improvements among them, instead of one round per distance left.
        // (The rounds' booked bytes already cover one read of every row.)
        done += sweep_to_fixpoint(g, n, m2, off, col, key, kOneHop, b, false, false, flag_base, false);
        break;
      }
    }
    if (flag_base + batch + 2 >= kMaxFlags) {
      fill_u32(reinterpret_cast<uint32_t *>(g->flags), kMaxFlags, 0, g->stream);
      flag_base = 0;
    }
    int slots[16];
    for (uint32_t i = 0; i < batch; ++i) {
      const uint32_t f = flag_base + i;
      const int *prev = i == 0 ? nullptr : &g->flags[f - 1];
      slots[i] = g->timer.start(KC_K_ROUND);
      // Round 0 pushes from about a tenth of the rows; a hub gets about a
      // tenth of its degree in pushes, serialised on its key.  Once that
      // exceeds what the round's other pushes take (past about 1/512 of the
      // level's edge slots), the largest hubs pull first.  (Pulling round 0
      // everywhere, every row scanned against flags of the distance-0
      // vertices, was measured slower at R-MAT 24: 2.36 against 1.35 ms.)
      const uint64_t hub_pull_deg = t + i == 0 ? std::max<uint64_t>(kHubPullMin, m2 / 512) : 0;
      // Round 0 of the degree-ordered level 0 walks each bin in level-index
      // order: a target ends at the smallest index among its distance-0
      // neighbours, which then tends to land first, so that the later
      // pushes to it read a settled key and skip their atomic.  Round 0
      // dispatches the hub chunks first (there when their rows carry small
      // indices, on the other levels always: the rounds of an R-MAT 24 step
      // 1.14 -> 1.09 ms, RHG 2^24 unchanged).  (The result does not depend
      // on the order.)
      const Bins &rb = t + i == 0 && index_round0 ? round0_bins(g, b) : b;
      push_round(off, col, key, kOneHop, rb, t + i, &g->flags[f], prev, hub_pull_deg, g->stream,
                 t + i == 0 && (!index_round0 || g->l0i_hubs_first));
      g->timer.stop(slots[i], 0);
      if (first_slot < 0) first_slot = slots[i];
    }
    d2h(g, g->h_flags, &g->flags[flag_base], batch * sizeof(int));
    if (deciding) {  // edge slots at and above the next round's distance, for its direction
      round_volumes(n, off, key, t + batch, g->sizes + 12, g->stream);
      d2h(g, g->h_sizes + 12, g->sizes + 12, 2 * sizeof(uint64_t));
    }
    sync(g);
    uint32_t i = 0;
    while (i < batch && g->h_flags[i] != 0) ++i;
    const uint32_t worked = i < batch ? i + 1 : batch;
    done += worked;
    flag_base += batch;
    t += batch;
    if (i < batch) break;
  }
  // The rounds together do the work of one sweep: every row is read once,
  // when its vertex fires (offsets, neighbour ids, neighbour keys), and every
  // key is read and settled once: 8(n+1) + 12 2m + 16n bytes, booked on the
  // level's first round.
  if (first_slot >= 0) g->timer.retag(first_slot, KC_K_ROUND, 8 * (n + 1) + 12 * m2 + 16 * n);
  return done;
}

void ensure_work(kc_graph *g, uint64_t n, uint64_t m2);
void ensure_learned_buffers(kc_graph *g, uint64_t n = 0);
void ensure_pieces(kc_graph *g, const BinCfg &cfg, uint64_t split);

struct RunOut {
  uint64_t sweeps = 0, radius_sum = 0;
  uint32_t levels = 0;
  uint64_t replicas = 0;
  uint32_t hubs = 0;
};

// Level 0 with its hubs split into proxies (kc_options.replicate): the
// reference's DistributeHighDegreeVertices (exponential_contraction.h:935-1253)
// with every vertex its own PE, so that every edge of a hub is a remote one.
// A hub (degree >= hd_threshold * sqrt(n0), utils.h:108-117) of degree d gets
// floor(sqrt(d)) proxy slots; each of its edges goes to the slot replica_pick
// names, and an edge between two hubs joins the two proxies.  The slots that
// receive an edge become proxies, numbered n0, n0 + 1, ... in hub order, each
// joined to its hub; the hub's row lists its proxies.  Proxies take their
// hub's id (they draw its deviate, :483-503) and, with indices above every
// vertex, lose every tie.  Returns false when there is no hub.
struct Replicated {
  uint64_t n = 0, m2 = 0;
  uint64_t *off = nullptr;
  uint32_t *col = nullptr, *gid = nullptr;
  uint32_t H = 0, R = 0;
};

template <class T>
T *upload(kc_graph *g, const std::vector<T> &v) {
  T *d = g->arena.alloc_n<T>(v.size() + 1);
  if (!v.empty()) KC_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

bool build_replicated(kc_graph *g, const kc_options *o, uint64_t n0, uint64_t m2, const uint64_t *off,
                      const uint32_t *col, const uint32_t *gid, Replicated &out) {
  constexpr uint64_t kChunk = 16384;  // hub edges per block
  hipStream_t s = g->stream;
  if (!(o->hd_threshold >= 0.0)) fail(KC_ERR_INVALID, "hd_threshold must be >= 0");
  const uint64_t limit = std::max<uint64_t>(1, static_cast<uint64_t>(o->hd_threshold * std::sqrt(static_cast<double>(n0))));
  uint32_t *list = g->arena.alloc_n<uint32_t>(std::min<uint64_t>(n0, m2 / limit + 1) + 1);
  find_hubs(n0, off, limit, list, g->small + 6, s);
  uint32_t H = 0;
  KC_HIP(hipMemcpyAsync(&H, g->small + 6, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  sync(g);
  if (H == 0) return false;
  std::vector<uint32_t> hubs(H);
  KC_HIP(hipMemcpy(hubs.data(), list, 4ull * H, hipMemcpyDeviceToHost));
  std::sort(hubs.begin(), hubs.end());
  // the hubs' row bounds
  std::vector<uint64_t> rows(2ull * H);
  const uint32_t *d_hubs = upload(g, hubs);
  uint64_t *d_rows = g->arena.alloc_n<uint64_t>(2ull * H);
  row_bounds(H, d_hubs, off, d_rows, s);
  KC_HIP(hipMemcpyAsync(rows.data(), d_rows, 16ull * H, hipMemcpyDeviceToHost, s));
  sync(g);
  std::vector<uint32_t> k(H);
  std::vector<uint64_t> slot_base(H + 1, 0);
  std::vector<HubEdgeChunk> chunks;
  for (uint32_t h = 0; h < H; ++h) {
    const uint64_t d = rows[2ull * h + 1] - rows[2ull * h];
    // min(P - 1, sqrt(d)) proxies (:1040) with P = n0 PEs
    k[h] = static_cast<uint32_t>(std::min<uint64_t>(n0 - 1, static_cast<uint64_t>(std::sqrt(static_cast<double>(d)))));
    if (k[h] == 0) k[h] = 1;
    slot_base[h + 1] = slot_base[h] + k[h];
    for (uint64_t e = rows[2ull * h]; e < rows[2ull * h + 1]; e += kChunk)
      chunks.push_back({h, 0, e, std::min(rows[2ull * h + 1], e + kChunk)});
  }
  const uint64_t S = slot_base[H];
  ReplicaArgs r{};
  r.off = off;
  r.col = col;
  r.gid = gid;
  r.seed = o->seed;
  r.hubs = d_hubs;
  r.k = upload(g, k);
  r.slot_base = upload(g, slot_base);
  const HubEdgeChunk *d_chunks = upload(g, chunks);
  r.hubidx = g->arena.alloc_n<uint32_t>(n0);
  r.cnt = g->arena.alloc_n<uint32_t>(S);
  fill_u32(r.cnt, S, 0, s);
  replica_count(r, n0, H, d_chunks, static_cast<uint32_t>(chunks.size()), s);
  std::vector<uint32_t> cnt(S);
  KC_HIP(hipMemcpyAsync(cnt.data(), r.cnt, 4 * S, hipMemcpyDeviceToHost, s));
  sync(g);
  // the slots with an edge become proxies, in slot (= hub, then slot) order
  std::vector<uint32_t> rep_of_slot(S, kNone), nexist(H, 0), first_rep(H, 0), slot_of_rep, root_of_rep;
  for (uint32_t h = 0; h < H; ++h) {
    first_rep[h] = static_cast<uint32_t>(n0 + slot_of_rep.size());
    for (uint64_t sl = slot_base[h]; sl < slot_base[h + 1]; ++sl)
      if (cnt[sl]) {
        rep_of_slot[sl] = static_cast<uint32_t>(n0 + slot_of_rep.size());
        slot_of_rep.push_back(static_cast<uint32_t>(sl));
        root_of_rep.push_back(hubs[h]);
        ++nexist[h];
      }
  }
  const uint64_t R = slot_of_rep.size();
  if (n0 + R >= 0xFFFFFFFFull || m2 + 2 * R >= (1ull << 32))
    fail(KC_ERR_INVALID, "replicated level 0 exceeds 32-bit ids (%llu proxies)", static_cast<unsigned long long>(R));
  r.rep_of_slot = upload(g, rep_of_slot);
  r.nexist = upload(g, nexist);
  r.first_rep = upload(g, first_rep);
  r.slot_of_rep = upload(g, slot_of_rep);
  r.root_of_rep = upload(g, root_of_rep);
  out.n = n0 + R;
  out.m2 = m2 + 2 * R;
  out.H = H;
  out.R = static_cast<uint32_t>(R);
  ensure_work(g, out.n, out.m2);
  uint32_t *len = g->arena.alloc_n<uint32_t>(out.n);
  out.off = g->arena.alloc_n<uint64_t>(out.n + 1);
  out.col = g->arena.alloc_n<uint32_t>(out.m2);
  out.gid = g->arena.alloc_n<uint32_t>(out.n);
  replica_lengths(r, n0, static_cast<uint32_t>(R), len, s);
  csr_offsets(out.n, len, out.off, g->sizes, g->scan_scratch, s);
  uint32_t *cursor = g->arena.alloc_n<uint32_t>(S);
  fill_u32(cursor, S, 0, s);
  replica_fill(r, n0, H, static_cast<uint32_t>(R), d_chunks, static_cast<uint32_t>(chunks.size()), out.off, cursor,
               out.col, out.gid, s);
  d2h(g, g->h_sizes + 2, g->sizes + 2, sizeof(uint64_t));
  sync(g);
  if (g->h_sizes[2] != out.m2)
    fail(KC_ERR_INVALID, "internal: replicated level 0 has %llu slots, expected %llu",
         static_cast<unsigned long long>(g->h_sizes[2]), static_cast<unsigned long long>(out.m2));
  return true;
}

// Exponential (or local) contraction down to an edgeless graph or the base case.
void run_contraction(kc_graph *g, const kc_options *o, RunOut &out) {
  const bool local = o->variant == KC_VARIANT_LOCAL;
  const BinCfg cfg = bin_cfg(g, o);
  // degree-0 vertices are their own components and take no part in the
  // contraction (the reference sets them inactive at load, graph_io.h:380)
  identity_u32(g->n, g->labels, g->stream);
  fill_u32(g->small + 5, 1, 0, g->stream);
  uint64_t n_l = g->n0, m2_l = g->m2;
  const uint64_t *off_l = g->off0c;
  const uint32_t *col_l = g->col0c;
  uint32_t *gid_l = g->gid0;
  // Level 0 in degree order when built (and the level is not already the base
  // case, whose pointer jumping indexes keys by level index): keys and rows by
  // storage position, key labels and gid_l by level index as before.
  // (the pre-contraction's blocks are ranges of level indices: input order)
  const bool ordered = g->idx0 && !o->input_order && o->pre_contract == KC_PRE_NONE && !o->replicate &&
                       !(o->seq_limit > 0 && m2_l / 2 <= o->seq_limit) && m2_l > 0;
  int pre_left = o->pre_contract == KC_PRE_DOUBLE ? 2 : o->pre_contract != KC_PRE_NONE ? 1 : 0;
  const uint32_t *idx_l = nullptr, *pos_l = nullptr;  // storage maps of the current level (null: identity)
  const uint32_t *dev_gid_l = gid_l;                   // input id of each storage position
  if (ordered) {
    off_l = g->off0p;
    col_l = g->col0p;
    idx_l = g->idx0;
    pos_l = g->pos0;
    dev_gid_l = g->gidp0;
  }
  const uint32_t *canon_gid = ordered ? g->gidp0 : g->gid0;  // input id of each level-0 storage position
  bool input_level0 = true;  // the level is the input's level 0 (not replicated, not contracted)
  if (o->replicate && !local && m2_l > 0) {
    // level 0 with its hubs split into proxies (input order); the canonical
    // labels below read the first n0 positions, the real vertices
    Replicated rp;
    if (build_replicated(g, o, n_l, m2_l, off_l, col_l, gid_l, rp)) {
      n_l = rp.n;
      m2_l = rp.m2;
      off_l = rp.off;
      col_l = rp.col;
      gid_l = rp.gid;
      dev_gid_l = rp.gid;
      input_level0 = false;
      out.hubs = rp.H;
      out.replicas = rp.R;
      if (o->learned || g->collect) ensure_learned_buffers(g, n_l);
    }
  }
  uint32_t iteration = 1;
  uint32_t flag_base = 0;
  int pp = 0;
  fill_u32(reinterpret_cast<uint32_t *>(g->flags), kMaxFlags, 0, g->stream);
  g->levels.clear();
  g->base_case_used = false;
  g->base_case_rounds = 0;
  const int max_levels = o->max_levels > 0 ? o->max_levels : 4096;
  DeviateArgs da{o->seed, 0, g->non_isolated, o->deviate_fma, o->deviate_stream, g->d_table, g->table_k};
  bool top_from_keys = false;  // top rep computed by the base case into rep of that level
  uint32_t *top_rep = nullptr;
  Bins b;
  bool bins_valid = false;  // bins of the current level graph (kept across stall levels)
  // a contracted level whose first kLevelHotRows vertices hold at least
  // kHotShareFactor times their proportional share of its edge slots: its
  // relabel keeps their labels in LDS (R-MAT 24 level 1, whose low indices
  // are the hub clusters; not the grids)
  bool level_hot = false;

  while (n_l > 0) {
    // a pre-contraction level comes before the cut-off, as the local phase
    // comes before the decomposition
    const bool pre = pre_left > 0 && m2_l > 0;
    if (pre) --pre_left;
    if (!pre && (m2_l == 0 || (o->seq_limit > 0 && m2_l / 2 <= o->seq_limit))) {
      // Base case (RunSequentialCC, exponential_contraction.h:908-922): min-label
      // propagation with pointer jumping on the remaining graph.
      if (!bins_valid) b = make_bins(g, n_l, off_l, cfg);
      {
        TimedScope ts(g->timer, KC_K_DEVIATES, 8 * n_l);
        init_keys_index(n_l, g->key, g->stream);
      }
      uint64_t r = 0;
      if (m2_l > 0) r = sweep_to_fixpoint(g, n_l, m2_l, off_l, col_l, g->key, 0, b, false, true, flag_base);
      g->base_case_used = m2_l > 0;
      g->base_case_rounds = r;
      out.sweeps += r;
      top_rep = g->rep_a;
      {
        TimedScope ts(g->timer, KC_K_UNDO, 16 * n_l);
        rep_from_keys(n_l, g->key, gid_l, top_rep, g->stream);
      }
      top_from_keys = true;
      break;
    }
    if (static_cast<int>(g->levels.size()) >= max_levels)
      fail(KC_ERR_INVALID, "contraction did not finish within %d levels", max_levels);

    if (!bins_valid) {
      b = make_bins(g, n_l, off_l, cfg);
      bins_valid = true;
    }
    uint64_t sweeps = 0;
    if (pre) {
      // components inside the blocks: min-label propagation over the edges
      // inside a block, with pointer jumping, to the fixpoint
      const uint64_t bsize = ceil_div(n_l, static_cast<uint64_t>(o->pre_blocks));
      {
        TimedScope ts(g->timer, KC_K_DEVIATES, 8 * n_l);
        init_keys_index(n_l, g->key, g->stream);
      }
      // (sweeps over the edges inside the blocks, with pointer jumping, in batches)
      sweeps = sweep_to_fixpoint(g, n_l, m2_l, off_l, col_l, g->key, 0, b, false, true, flag_base, true, bsize);
      if (o->pre_contract == KC_PRE_LOCAL) {
        TimedScope ts(g->timer, KC_K_SWEEP, 8 * (n_l + 1) + 4 * m2_l + 8 * n_l);
        block_interface(off_l, col_l, n_l, bsize, g->key, g->stream);
      }
    } else {
      // deviates -> keys
      LearnedArgs la{nullptr};
      if (!local && o->learned) {
        // scores of this level's vertices delay their deviates; the input's
        // level 0 reuses its scores from an earlier call with the same model
        const bool l0 = input_level0 && iteration == 1;
        if (l0 && g->l0_score && g->l0_score_mlp == g->mlp && g->l0_score_version == g->mlp->version &&
            g->l0_score_off == off_l) {
          TimedScope lt(g->timer, KC_K_LEARNED, 4 * n_l + n_l);
          score_delays(n_l, g->l0_score, o->learned_alpha, g->delay, g->stream);
        } else {
          TimedScope lt(g->timer, KC_K_LEARNED, 8 * (n_l + 1) + 4 * n_l + 40 * n_l + 24 * n_l + n_l);
          level_degrees(n_l, off_l, g->ldeg, g->stream);
          level_features(n_l, off_l, col_l, g->ldeg, iteration - 1, g->feat, g->stream);
          mlp_delays(g->mlp, n_l, g->feat, g->target, o->learned_alpha, g->delay, g->stream);
          if (l0) {
            if (!g->l0_score) g->l0_score = g->dalloc<float>(g->n0);
            KC_HIP(hipMemcpyAsync(g->l0_score, g->target, n_l * sizeof(float), hipMemcpyDeviceToDevice, g->stream));
            g->l0_score_mlp = g->mlp;
            g->l0_score_version = g->mlp->version;
            g->l0_score_off = off_l;
          }
        }
        la.delay = g->delay;
      }
      {
        TimedScope ts(g->timer, KC_K_DEVIATES, 8 * n_l + (gid_l ? 4 * n_l : 0) + (la.delay ? n_l : 0));
        if (local) {
          init_keys_uniform(n_l, dev_gid_l, idx_l, o->seed, iteration, g->key, g->stream);
        } else {
          da.iteration = iteration;
          init_keys_exp(n_l, dev_gid_l, idx_l, da, la, g->key, g->stream);
        }
      }
      if (local) {
        // two synchronous min passes over the closed neighbourhood
        // (local_contraction.h:195-274), then a counting pass on the final keys
        for (int pass = 0; pass < 2; ++pass) {
          uint64_t *kin = pass == 0 ? g->key : g->key2;
          uint64_t *kout = pass == 0 ? g->key2 : g->key;
          KC_HIP(hipMemcpyAsync(kout, kin, n_l * 8, hipMemcpyDeviceToDevice, g->stream));
          TimedScope ts(g->timer, KC_K_SWEEP, sweep_bytes(n_l, m2_l, false));
          sweep(kSweepJacobi, false, off_l, col_l, kin, kout, 0, b, g->deg_inter, g->hub_inter,
                &g->flags[kMaxFlags - 1], nullptr, g->stream);
        }
        sweeps = 2;
      } else if (o->use_bfs) {
        const bool choose = o->pull_rounds > 0 || (o->pull_rounds == 0 && n_l >= kPullRoundsMin &&
                                                   m2_l >= kDirectionMinDegree * n_l);
        sweeps = rounds_to_fixpoint(g, n_l, m2_l, off_l, col_l, g->key, b, cfg, flag_base,
                                    o->tail_sweeps < 0 ? -1 : o->tail_sweeps == 0 ? kTailShift : o->tail_sweeps,
                                    choose, pos_l != nullptr && off_l == g->off0p && g->l0_hot_skewed);
      } else {
        // Levels below 2^24 vertices whose key distances stay below 255 sweep
        // packed 32-bit keys (pack_keys; the buffer of the cluster ranks is
        // free until they are built).  A distance only falls, so a key plus
        // one hop stays within 32 bits.
        uint32_t *pk = n_l >= kPackedSweepsMin && n_l <= (1ull << kPackedLabelBits) ? g->crank : nullptr;
        if (pk) {
          pack_keys(n_l, g->key, pk, g->small + 7, g->stream);
          d2h(g, g->h_small + 7, g->small + 7, sizeof(uint32_t));
          sync(g);
          if (g->h_small[7]) pk = nullptr;
        }
        sweeps = sweep_to_fixpoint(g, n_l, m2_l, off_l, col_l, g->key, kOneHop, b, false, false, flag_base, true, 0,
                                   pk);
        if (pk) unpack_keys(n_l, pk, g->key, g->stream);
      }
    }
    if (!pre && !local && !o->use_bfs && sweeps == 1) {
      // Stall: the first sweep changed no key, so every vertex is its own
      // centre and keeps an edge (no level holds an isolated vertex); the
      // contracted graph equals this one.  The level is recorded and the next
      // iteration draws new deviates on the same graph.
      LevelRec rec;
      rec.stall = true;
      rec.idx = idx_l;
      rec.n = n_l;
      rec.m2 = m2_l;
      rec.gid = gid_l;
      rec.sweeps = 1;
      if (o->record_trace) {
        rec.keys_copy = g->arena.alloc_n<uint64_t>(n_l);
        KC_HIP(hipMemcpyAsync(rec.keys_copy, g->key, n_l * 8, hipMemcpyDeviceToDevice, g->stream));
      }
      out.sweeps += 1;
      g->levels.push_back(rec);
      ++iteration;
      continue;
    }
    LevelRec rec;
    rec.idx = idx_l;
    rec.n = n_l;
    rec.m2 = m2_l;
    rec.gid = gid_l;
    rec.sweeps = pre ? 0 : sweeps;  // a pre-contraction level has no decomposition sweeps
    rec.lab = g->arena.alloc_n<uint32_t>(n_l);
    rec.newidx = g->arena.alloc_n<uint32_t>(n_l);
    uint32_t *gid_next = g->arena.alloc_n<uint32_t>(n_l);
    if (o->record_trace) {
      rec.keys_copy = g->arena.alloc_n<uint64_t>(n_l);
      KC_HIP(hipMemcpyAsync(rec.keys_copy, g->key, n_l * 8, hipMemcpyDeviceToDevice, g->stream));
    }
    fill_u32(g->small + 4, 1, 0, g->stream);
    // dense cluster ids in label order: the buckets below hold them, so the
    // dedup works over the C clusters instead of the n_l labels; the labels'
    // pass marks the used ones
    fill_u32(g->cid, n_l, 0, g->stream);  // the used-label flags, then the ids
    {
      TimedScope ts(g->timer, KC_K_COUNT, 8 * n_l + 4 * n_l + 4 * n_l + (local ? 0 : 8 * n_l));
      labels_radius(n_l, g->key, pos_l, rec.lab, g->small + 4, !local && !pre, g->stream, g->cid);
    }
    if (g->collect && !pre) {
      // trace of the plain method: features and the log2 size of each vertex's cluster
      level_degrees(n_l, off_l, g->ldeg, g->stream);
      level_features(n_l, off_l, col_l, g->ldeg, iteration - 1, g->feat, g->stream);
      cluster_targets(n_l, rec.lab, g->csize, g->target, g->stream);
      mlp_append(g->collect, n_l, g->feat, g->target, g->collect_stride, g->stream);
    }
    // on a degree-ordered level 0 whose first kHotClusters positions hold a
    // large share of the edge slots, the clusters of those positions take the
    // ids below 65536, so the relabel keeps those positions' ids in LDS at 16
    // bits (twice as many as at 32)
    const bool hot_ids = pos_l != nullptr && g->l0_hot_skewed;
    {
      TimedScope ts(g->timer, KC_K_COUNT, 8 * n_l);
      if (hot_ids)
        cluster_ranks_hot(n_l, std::min<uint64_t>(n_l, kHotClusters), rec.lab, g->cid, g->crank, g->sizes + 9,
                          g->scan_scratch, g->stream);
      else
        cluster_ranks(n_l, g->cid, g->crank, g->sizes + 9, g->scan_scratch, g->stream);
    }
    {
      TimedScope ts(g->timer, KC_K_COUNT, 12 * n_l);
      cluster_ids(n_l, rec.lab, g->crank, g->cid, g->stream);
    }
    // relabel: every row packs the cluster ids of its far ends outside its
    // cluster at the start of its slots; the slot buffer later takes the unique
    // runs of the dedup.  Its bytes (offsets, neighbour ids and their ids, the
    // row's id and count, the packed writes) are booked once P is known.
    uint32_t *relab = g->slots;
    // the degree-ordered level 0 relabels the rows above the 16-lane bin in
    // owner pieces, which then stand in for its wave rows and hub chunks in
    // the count fold and the bucket copy
    const bool split = hot_ids && off_l == g->off0p && n_l == g->n0 && o->relabel_split >= 0;
    Bins rb = b;
    uint32_t *r_inter = g->hub_inter, *r_pos = g->hub_pos;
    uint64_t *r_prefix = g->hub_prefix;
    if (split) {
      ensure_pieces(g, cfg, o->relabel_split > 1 ? static_cast<uint64_t>(o->relabel_split) : kSplitRowsAbove);
      rb.listW = nullptr;
      rb.nW = 0;
      rb.hubs = g->l0_pieces;
      rb.nH = static_cast<uint32_t>(g->l0_npieces);
      finish_bins(rb);
      r_inter = g->l0_piece_inter;
      r_pos = g->l0_piece_pos;
      r_prefix = g->l0_piece_prefix;
    }
    int relabel_slot;
    {
      TimedScope ts(g->timer, KC_K_RELABEL, 0);
      relabel_slot = ts.slot;
      // the hot-label launch pays off on the degree-ordered level 0
      if (split)
        relabel_split(off_l, col_l, g->cid, n_l, b, g->l0_pieces, g->l0_part_list, g->l0_part_at, relab,
                      g->deg_inter, r_inter, n_l, g->stream);
      else
        relabel(off_l, col_l, g->cid, b, relab, g->deg_inter, g->hub_inter, pos_l || level_hot ? n_l : 0,
                g->stream, hot_ids);
    }
    if (rb.nH) {
      TimedScope ts(g->timer, KC_K_COUNT, 12ull * rb.nH + 4ull * rb.nH);
      hub_fold(rb, r_inter, g->deg_inter, r_pos, r_prefix, g->scan_scratch, g->stream);
    }
    // clusters as segments of the (label, vertex) list sorted by label
    {
      TimedScope ts(g->timer, KC_K_SCAN, 2 * 4 * n_l + 8 * n_l + 8 * n_l);
      select_inter(n_l, g->key, g->deg_inter, g->sk, g->sv, g->sizes + 8, g->scan_scratch, g->stream);
    }
    d2h(g, g->h_sizes + 8, g->sizes + 8, 2 * sizeof(uint64_t));
    sync(g);
    const uint64_t N = g->h_sizes[8];
    const uint64_t clusters = g->h_sizes[9];
    int bits = 1;
    while (bits < 32 && (1ull << bits) < n_l) ++bits;
    int which = 0;
    {
      TimedScope ts(g->timer, KC_K_SORT, 24 * N * static_cast<uint64_t>((bits + 7) / 8));
      which = radix_sort_pairs(N, bits, g->sk, g->sv, g->sk2, g->sv2, g->sort_scratch, g->stream);
    }
    const uint32_t *sk = which ? g->sk2 : g->sk, *sv = which ? g->sv2 : g->sv;
    fill_u32(rec.newidx, n_l, kNone, g->stream);
    {
      TimedScope ts(g->timer, KC_K_SCAN, 2 * (8 * N + 4 * N) + 4 * N + 20 * n_l / 8);
      // the sort's other buffer pair is free: it caches the gathered counts
      segments(N, sk, sv, g->deg_inter, gid_l, g->crank, g->pos_of, rec.newidx, g->seg_of_cid, gid_next, g->seg_off,
               g->sizes, which ? g->sk : g->sk2, g->scan_scratch, g->stream);
    }
    // medium and huge buckets listed before the sizes come back (the
    // clusters bound n_next)
    {
      TimedScope ts(g->timer, KC_K_SCAN, 16 * clusters + 4 * clusters / 8);
      classify_buckets(std::min(N, clusters), g->sizes, g->seg_off, g->dl_medium, g->dl_huge, g->sizes + 3,
                       g->scan_scratch, g->stream);
    }
    d2h(g, g->h_sizes, g->sizes, 5 * sizeof(uint64_t));
    d2h(g, g->h_small + 4, g->small + 4, sizeof(uint32_t));
    sync(g);
    const uint64_t n_next = N ? g->h_sizes[0] : 0;
    const uint64_t P = N ? g->h_sizes[1] : 0;
    const uint64_t n_medium = N ? g->h_sizes[3] : 0, n_huge = N ? g->h_sizes[4] : 0;
    rec.radius = local || pre ? 0 : g->h_small[4];
    if (!pre) out.sweeps += sweeps;  // the block passes are not decomposition sweeps
    out.radius_sum += rec.radius;
    g->levels.push_back(rec);
    if (n_next == 0) break;  // every cluster is a whole component

    // gather: every row copies its cluster edges to its position in its
    // cluster's bucket, which lives in the next level's column buffer until
    // the columns are built
    uint64_t *off_next = g->off_pp[pp];
    uint32_t *col_next = g->col_pp[pp];
    uint32_t *bucket = col_next;
    {
      g->timer.retag(relabel_slot, KC_K_RELABEL, 8 * (n_l + 1) + 8 * m2_l + 8 * n_l + 4 * P);
      TimedScope ts(g->timer, KC_K_SCATTER, 8 * (n_l + 1) + 8 * n_l + 8 * P);
      scatter(off_l, relab, rb, g->deg_inter, r_inter, r_pos, g->pos_of, bucket, g->stream);
    }
    const int dedup_slot = dedup_buckets(g, n_next, g->seg_off, bucket, g->slots, g->uniq, g->dl_medium, g->dl_huge,
                                         n_medium, n_huge, P, clusters, g->scan_scratch, o->dedup_huge);
    // next-level CSR
    {
      TimedScope ts(g->timer, KC_K_SCAN, 4 * n_next + 8 * (n_next + 1));
      csr_offsets(n_next, g->uniq, off_next, g->sizes, g->scan_scratch, g->stream);
    }
    d2h(g, g->h_sizes + 2, g->sizes + 2, sizeof(uint64_t));
    const uint64_t hot_rows = std::min<uint64_t>(n_next, kLevelHotRows);
    d2h(g, g->h_sizes + 15, off_next + hot_rows, sizeof(uint64_t));
    sync(g);
    const uint64_t m2_next = g->h_sizes[2];
    level_hot = m2_next > 0 && static_cast<double>(g->h_sizes[15]) * n_next >=
                                   kHotShareFactor * static_cast<double>(hot_rows) * m2_next;
    // dedup: every bucket slot read once, the unique ones written, per
    // cluster its bucket bounds and its unique count
    g->timer.retag(dedup_slot, KC_K_DEDUP, 4 * P + 4 * m2_next + 8 * (n_next + 1) + 4 * n_next);
    {
      // unique runs read, mapped to next-level indices, written
      TimedScope ts(g->timer, KC_K_CSR, 4 * m2_next + 4 * m2_next + 4 * m2_next);
      gather_uniques(n_next, off_next, g->seg_off, g->slots, g->seg_of_cid, g->dl_medium, n_medium, g->dl_huge, n_huge,
                     g->row_cstart, g->scan_scratch, col_next, g->stream);
    }
    n_l = n_next;
    m2_l = m2_next;
    off_l = off_next;
    col_l = col_next;
    gid_l = gid_next;
    dev_gid_l = gid_next;
    idx_l = nullptr;
    pos_l = nullptr;
    pp ^= 1;
    bins_valid = false;
    input_level0 = false;
    if (!pre) ++iteration;
  }
  out.levels = static_cast<uint32_t>(g->levels.size());

  // Undo: top-down, rep_l[v] = rep_{l+1}[newidx[lab[v]]] or gid[lab[v]].
  uint32_t *rep_up = top_from_keys ? top_rep : nullptr;
  for (int l = static_cast<int>(g->levels.size()) - 1; l >= 0; --l) {
    const LevelRec &r = g->levels[l];
    if (r.stall) continue;  // identity contraction
    uint32_t *rep_out = (rep_up == g->rep_a) ? g->rep_b : g->rep_a;
    TimedScope ts(g->timer, KC_K_UNDO, 12 * r.n);
    undo(r.n, r.lab, r.newidx, r.gid, rep_up ? rep_up : g->rep_b, rep_out, g->stream);
    rep_up = rep_out;
  }
  if (!rep_up) {  // no level at all (n == 0)
    return;
  }
  // canonical labels: minimum vertex id per component
  const uint64_t n0 = g->n0;
  const uint64_t gid_bytes = canon_gid ? 4 * n0 : 0;
  fill_u32(g->minid, g->n, kNone, g->stream);
  {
    TimedScope ts(g->timer, KC_K_CANON, 8 * n0 + gid_bytes);
    canon_min(n0, rep_up, canon_gid, g->minid, g->stream);
  }
  {
    TimedScope ts(g->timer, KC_K_CANON, 12 * n0 + gid_bytes);
    canon_label(n0, rep_up, canon_gid, g->minid, g->labels, g->small + 5, reinterpret_cast<uint32_t *>(g->scan_scratch),
                g->stream);
  }
}

// Buffers of the learned centre proposals for levels of up to n vertices
// (level 0's size unless a larger replicated level 0 needs more).
void ensure_learned_buffers(kc_graph *g, uint64_t n) {
  n = std::max(n, g->n0);
  if (g->feat && n <= g->learned_cap) return;
  g->drealloc(g->feat, n * kMlpIn);
  g->drealloc(g->target, n);
  g->drealloc(g->ldeg, n);
  g->drealloc(g->csize, n);
  g->drealloc(g->delay, n);
  g->learned_cap = n;
}

// Min-label propagation (labels) or propagation + pointer jumping (shortcuts).
void run_propagation(kc_graph *g, const kc_options *o, RunOut &out) {
  const BinCfg cfg = bin_cfg(g, o);
  uint32_t flag_base = 0;
  fill_u32(reinterpret_cast<uint32_t *>(g->flags), kMaxFlags, 0, g->stream);
  g->levels.clear();
  Bins b = make_bins(g, g->n, g->off0, cfg);
  {
    TimedScope ts(g->timer, KC_K_DEVIATES, 8 * g->n);
    init_keys_index(g->n, g->key, g->stream);
  }
  out.sweeps = sweep_to_fixpoint(g, g->n, g->m2, g->off0, g->col0, g->key, 0, b, false,
                                 o->variant == KC_VARIANT_SHORTCUT, flag_base);
  out.levels = 0;
  TimedScope ts(g->timer, KC_K_CANON, 12 * g->n);
  labels_from_keys(g->n, g->key, g->labels, g->small + 5, reinterpret_cast<uint32_t *>(g->scan_scratch), g->stream);
}

}  // namespace

// ===========================================================================
// C ABI.
extern "C" {

const char *kc_last_error(void) { return g_last_error.c_str(); }

void kc_options_init(kc_options *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->variant = KC_VARIANT_EXP;
  o->seed = 1;
  o->beta = 0.1;
  o->seq_limit = 4096;
  o->deviate_fma = 1;
  o->deviate_stream = KC_STREAM_REFERENCE;
  o->max_levels = 4096;
  o->learned_alpha = 1.0f;
  o->hd_threshold = 16.0;
}

}  // extern "C"

namespace {

// A handle with its stream and events; the CSR and work buffers come later.
std::unique_ptr<kc_graph> new_handle() {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) fail(KC_ERR_HIP, "no HIP device available");
  std::unique_ptr<kc_graph> g(new kc_graph());
  KC_HIP(hipGetDevice(&g->device));
  KC_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  KC_HIP(hipEventCreate(&g->ev_begin));
  KC_HIP(hipEventCreate(&g->ev_end));
  KC_HIP(hipHostMalloc(reinterpret_cast<void **>(&g->h_sizes), 16 * sizeof(uint64_t), hipHostMallocDefault));
  KC_HIP(hipHostMalloc(reinterpret_cast<void **>(&g->h_small), 16 * sizeof(uint32_t), hipHostMallocDefault));
  KC_HIP(hipHostMalloc(reinterpret_cast<void **>(&g->h_flags), 64 * sizeof(int), hipHostMallocDefault));
  g->small = g->dalloc<uint32_t>(16);
  g->sizes = g->dalloc<uint64_t>(16);
  return g;
}

// Temporary device buffer freed at scope exit.
struct DevTemp {
  void *p = nullptr;
  explicit DevTemp(size_t bytes) {
    if (hipMalloc(&p, bytes ? bytes : 64) != hipSuccess) fail(KC_ERR_OOM, "hipMalloc of %zu bytes failed", bytes);
  }
  ~DevTemp() { (void)hipFree(p); }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// The level-sized work set for levels of up to cap_n vertices and cap_m2 edge
// slots (the previous set, if any, is freed first).
void alloc_work(kc_graph *g, uint64_t cap_n, uint64_t cap_m2) {
  g->free_work();
  const uint64_t n = cap_n, m2 = cap_m2;
  g->cap_n = n;
  g->cap_m2 = m2;
  g->key = g->walloc<uint64_t>(n);
  g->key2 = g->walloc<uint64_t>(n);
  g->deg_inter = g->walloc<uint32_t>(n);
  g->lists = g->walloc<uint32_t>(3 * n);
  g->hubs_cap = hub_chunk_cap(m2);
  g->hubs = g->walloc<HubChunk>(g->hubs_cap);
  g->flists = g->walloc<uint32_t>(3 * n);
  g->fhubs = g->walloc<HubChunk>(g->hubs_cap);
  g->hub_inter = g->walloc<uint32_t>(g->hubs_cap);
  g->slots = g->walloc<uint32_t>(m2);
  g->seg_of_cid = g->walloc<uint32_t>(n);
  g->crank = g->walloc<uint32_t>(n);
  g->cid = g->walloc<uint32_t>(n);
  g->uniq = g->walloc<uint32_t>(n);
  g->pos_of = g->walloc<uint32_t>(n);
  g->hub_pos = g->walloc<uint32_t>(g->hubs_cap);
  g->hub_prefix = g->walloc<uint64_t>(g->hubs_cap + 1);
  g->sk = g->walloc<uint32_t>(n);
  g->sv = g->walloc<uint32_t>(n);
  g->sk2 = g->walloc<uint32_t>(n);
  g->sv2 = g->walloc<uint32_t>(n);
  g->seg_off = g->walloc<uint64_t>(n + 1);
  g->sort_scratch = g->walloc<uint64_t>(radix_sort_scratch_words(n));
  g->rep_a = g->walloc<uint32_t>(n);
  g->rep_b = g->walloc<uint32_t>(n);
  g->scan_scratch = g->walloc<uint64_t>(std::max(scan_scratch_words(n + m2) + 1024, bin_scratch_words(n)));
  for (int i = 0; i < 2; ++i) {
    g->off_pp[i] = g->walloc<uint64_t>(n + 1);
    g->col_pp[i] = g->walloc<uint32_t>(m2);
  }
  g->dl_medium = g->walloc<uint32_t>(n);
  g->dl_huge = g->walloc<uint32_t>(n);
  g->row_cstart = g->walloc<uint64_t>(n + 2);
  g->hcount = g->walloc<uint32_t>(n);
  // level 0's cached bins live in the work set too
  g->l0_lists = nullptr;
  g->l0_hubs = nullptr;
  g->l0_off = nullptr;
  g->l0i_lists = nullptr;
  g->l0i_hubs = nullptr;
  g->l0i_valid = false;
}

Bins round0_bins(kc_graph *g, const Bins &b) {
  if (g->l0i_valid) return g->l0i_bins;
  if (!g->l0i_lists) {
    g->l0i_lists = g->walloc<uint32_t>(3 * g->cap_n);
    g->l0i_hubs = g->walloc<HubChunk>(g->hubs_cap);
  }
  Bins out = b;
  out.listT = g->l0i_lists;
  out.listS = g->l0i_lists + g->cap_n;
  out.listW = g->l0i_lists + 2 * g->cap_n;
  out.hubs = g->l0i_hubs;
  const uint64_t c = std::max<uint64_t>(std::max<uint64_t>(b.nT, b.nS), std::max<uint64_t>(b.nW, b.nH));
  {
    DevTemp k0(c * 4 + 64), v0(c * 4 + 64), k1(c * 4 + 64), v1(c * 4 + 64);
    DevTemp scratch(radix_sort_scratch_words(c) * 8 + 64);
    const uint32_t median = order_bins_by_index(b, g->idx0, g->n0, out, k0.as<uint32_t>(), v0.as<uint32_t>(),
                                                k1.as<uint32_t>(), v1.as<uint32_t>(), scratch.as<uint64_t>(), g->stream);
    sync(g);
    // a target ends at its smallest distance-0 label: when the hubs' labels
    // are small (R-MAT's hubs have the lowest ids), their pushes go first
    g->l0i_hubs_first = median != kNone && static_cast<uint64_t>(median) < g->n0 / 4;
  }
  g->l0i_bins = out;
  g->l0i_valid = true;
  return out;
}

// The pieces of the degree-ordered level 0 for the bins cfg, the rows above
// `split` slots cut by owner part (kept while the mid and large thresholds and
// the bound stay the same).
void ensure_pieces(kc_graph *g, const BinCfg &cfg, uint64_t split) {
  if (g->l0_pieces_valid && g->l0_piece_cfg.smax == cfg.smax && g->l0_piece_cfg.wmax == cfg.wmax &&
      g->l0_piece_split == split)
    return;
  g->l0_pieces_valid = false;
  const uint64_t n = g->n0;
  DevTemp cnt(n * 4 + 64), start((n + 1) * 8 + 64);
  count_pieces(n, g->off0p, g->col0p, cfg.smax, cfg.wmax, split, cnt.as<uint32_t>(), start.as<uint64_t>(),
               g->scan_scratch, g->stream);
  uint64_t total = 0;
  KC_HIP(hipMemcpyAsync(&total, start.as<uint64_t>() + n, sizeof(total), hipMemcpyDeviceToHost, g->stream));
  sync(g);
  if (total >= (1ull << 32)) fail(KC_ERR_INVALID, "internal: %llu relabel pieces", (unsigned long long)total);
  g->drealloc(g->l0_pieces, total);
  g->drealloc(g->l0_part_list, total);
  g->drealloc(g->l0_piece_inter, total);
  g->drealloc(g->l0_piece_pos, total);
  g->drealloc(g->l0_piece_prefix, total + 1);
  std::vector<uint8_t> part(total);
  if (total) {
    DevTemp dpart(total + 64);
    write_pieces(n, g->off0p, g->col0p, cfg.smax, cfg.wmax, split, start.as<uint64_t>(), g->l0_pieces,
                 dpart.as<uint8_t>(), g->stream);
    KC_HIP(hipMemcpyAsync(part.data(), dpart.as<uint8_t>(), total, hipMemcpyDeviceToHost, g->stream));
    sync(g);
  }
  // piece indices grouped by part (the shared ones last), ascending inside a part
  uint32_t at[kParts + 2] = {};
  for (uint8_t p : part) {
    if (p > kParts) fail(KC_ERR_INVALID, "internal: relabel piece of part %u", p);
    ++at[p + 1];
  }
  for (uint32_t p = 0; p <= kParts; ++p) at[p + 1] += at[p];
  std::vector<uint32_t> list(total);
  uint32_t next[kParts + 1];
  for (uint32_t p = 0; p <= kParts; ++p) next[p] = at[p];
  for (uint64_t i = 0; i < total; ++i) list[next[part[i]]++] = static_cast<uint32_t>(i);
  if (total)
    KC_HIP(hipMemcpyAsync(g->l0_part_list, list.data(), total * 4, hipMemcpyHostToDevice, g->stream));
  sync(g);
  for (uint32_t p = 0; p <= kParts + 1; ++p) g->l0_part_at[p] = at[p];
  g->l0_npieces = total;
  g->l0_piece_cfg = cfg;
  g->l0_piece_split = split;
  g->l0_pieces_valid = true;
}

// Grows the work set to hold a level of n vertices and m2 edge slots.
void ensure_work(kc_graph *g, uint64_t n, uint64_t m2) {
  if (n <= g->cap_n && m2 <= g->cap_m2) return;
  alloc_work(g, std::max(n, g->cap_n), std::max(m2, g->cap_m2));
}

// Work buffers sized for level 0 (g->n, g->m2 set, off0/col0 in place).
void init_work(kc_graph *g) {
  const uint64_t n = g->n, m2 = g->m2;
  alloc_work(g, n, m2);
  g->minid = g->dalloc<uint32_t>(n);
  g->labels = g->dalloc<uint32_t>(n);
  g->flags = g->dalloc<int>(kMaxFlags);
  // level 0 without the degree-0 vertices (rep_a holds the id map meanwhile)
  if (g->non_isolated == n) {
    g->n0 = n;
    g->gid0 = nullptr;
    g->off0c = g->off0;
    g->col0c = g->col0;
  } else {
    g->gid0 = g->dalloc<uint32_t>(n);
    g->off0c = g->dalloc<uint64_t>(n + 1);
    g->col0c = g->dalloc<uint32_t>(m2);
    drop_isolated(n, g->off0, g->col0, m2, g->rep_a, g->gid0, g->off0c, g->col0c, g->sizes + 9, g->scan_scratch,
                  g->stream);
    d2h(g, g->h_sizes + 9, g->sizes + 9, sizeof(uint64_t));
    sync(g);
    g->n0 = g->h_sizes[9];
    if (g->n0 != g->non_isolated)
      fail(KC_ERR_INVALID, "internal: %llu vertices of degree > 0, expected %llu", (unsigned long long)g->n0,
           (unsigned long long)g->non_isolated);
  }
  if (g->n0 >= kOrderMinVertices) {
    const uint64_t n0 = g->n0;
    g->idx0 = g->dalloc<uint32_t>(n0);
    g->pos0 = g->dalloc<uint32_t>(n0);
    g->gidp0 = g->dalloc<uint32_t>(n0);
    g->off0p = g->dalloc<uint64_t>(n0 + 1);
    g->col0p = g->dalloc<uint32_t>(m2);
    level0_order(n0, g->off0c, g->col0c, g->gid0, g->idx0, g->pos0, g->gidp0, g->off0p, g->col0p, g->sk, g->sv, g->sk2,
                 g->sv2, g->sort_scratch, g->scan_scratch, g->stream);
    // every row grouped by the owner part of its targets, ascending inside a
    // group: the owner-split relabel cuts rows at the group bounds, and a
    // row's gathers walk up the (degree-ordered) storage, neighbouring entries
    // sharing cache lines more often
    {
      DevTemp k0(m2 * 4 + 64), v0(m2 * 4 + 64), k1(m2 * 4 + 64), v1(m2 * 4 + 64);
      DevTemp scratch(radix_sort_scratch_words(m2) * 8 + 64);
      sort_rows(n0, g->off0p, g->col0p, m2, k0.as<uint32_t>(), v0.as<uint32_t>(), k1.as<uint32_t>(), v1.as<uint32_t>(),
                scratch.as<uint64_t>(), g->stream);
      sync(g);
    }
    {
      const uint64_t k = std::min<uint64_t>(n0, kHotClusters);
      uint64_t slots_k = 0;
      KC_HIP(hipMemcpy(&slots_k, g->off0p + k, sizeof(slots_k), hipMemcpyDeviceToHost));
      g->l0_hot_skewed = m2 > 0 && static_cast<double>(slots_k) * n0 >= kHotShareFactor * static_cast<double>(k) * m2;
    }
  }
  KC_HIP(hipDeviceSynchronize());
}

// Symmetric, self-loop-free, duplicate-free CSR on the device from an edge
// list (rows deduplicated in LDS hash tables, so the order inside a row is
// not fixed).  Sets g->off0/col0/m2.
void build_device_csr(kc_graph *g, uint64_t n, const uint32_t *src, const uint32_t *dst, uint64_t count) {
  hipStream_t s = g->stream;
  DevTemp deg((n + 1) * 4), off_tmp((n + 2) * 8), fill((n + 1) * 4), uniq((n + 1) * 4);
  DevTemp scratch((scan_scratch_words(n + 2 * count) + 1024) * 8);
  fill_u32(deg.as<uint32_t>(), n, 0, s);
  deg_count(count, src, dst, deg.as<uint32_t>(), s);
  csr_offsets(n, deg.as<uint32_t>(), off_tmp.as<uint64_t>(), g->sizes, scratch.as<uint64_t>(), s);
  KC_HIP(hipMemcpyAsync(g->h_sizes + 2, g->sizes + 2, 8, hipMemcpyDeviceToHost, s));
  KC_HIP(hipStreamSynchronize(s));
  const uint64_t P = g->h_sizes[2];
  if (P >= (1ull << 32)) fail(KC_ERR_INVALID, "2m=%llu exceeds 2^32 directed edges", (unsigned long long)P);
  DevTemp bucket(P * 4 + 64);
  fill_u32(fill.as<uint32_t>(), n, 0, s);
  edge_scatter(count, src, dst, off_tmp.as<uint64_t>(), fill.as<uint32_t>(), bucket.as<uint32_t>(), s);
  DevTemp lists((2 * n + 8) * 4);
  uint32_t *medium = lists.as<uint32_t>();
  uint32_t *huge = medium + n + 2;
  DevTemp runs(P * 4 + 64);
  classify_buckets(n, nullptr, off_tmp.as<uint64_t>(), medium, huge, g->sizes + 3, scratch.as<uint64_t>(), s);
  KC_HIP(hipMemcpyAsync(g->h_sizes + 3, g->sizes + 3, 16, hipMemcpyDeviceToHost, s));
  KC_HIP(hipStreamSynchronize(s));
  const uint64_t n_medium = g->h_sizes[3], n_huge = g->h_sizes[4];
  dedup_buckets(g, n, off_tmp.as<uint64_t>(), bucket.as<uint32_t>(), runs.as<uint32_t>(), uniq.as<uint32_t>(), medium,
                huge, n_medium, n_huge, P, n, scratch.as<uint64_t>());
  g->off0 = g->dalloc<uint64_t>(n + 1);
  csr_offsets(n, uniq.as<uint32_t>(), g->off0, g->sizes, scratch.as<uint64_t>(), s);
  KC_HIP(hipMemcpyAsync(g->h_sizes + 2, g->sizes + 2, 8, hipMemcpyDeviceToHost, s));
  KC_HIP(hipStreamSynchronize(s));
  g->m2 = g->h_sizes[2];
  g->col0 = g->dalloc<uint32_t>(g->m2);
  DevTemp cstart((n_medium + n_huge + 2) * 8);
  gather_uniques(n, g->off0, off_tmp.as<uint64_t>(), runs.as<uint32_t>(), nullptr, medium, n_medium, huge, n_huge,
                 cstart.as<uint64_t>(), scratch.as<uint64_t>(), g->col0, s);
  unsigned long long *cnt = reinterpret_cast<unsigned long long *>(g->sizes + 6);
  fill_u64(g->sizes + 6, 1, 0, s);
  nonzero_rows(n, g->off0, cnt, s);
  KC_HIP(hipMemcpyAsync(g->h_sizes + 6, g->sizes + 6, 8, hipMemcpyDeviceToHost, s));
  KC_HIP(hipStreamSynchronize(s));
  g->non_isolated = g->h_sizes[6];
}

}  // namespace

extern "C" {

int kc_graph_create(const kc_csr_host *csr, kc_graph **out) {
  return guarded([&] {
    if (!csr || !out) fail(KC_ERR_INVALID, "null argument");
    *out = nullptr;
    const uint64_t n = csr->n;
    if (n >= 0xFFFFFFFFull) fail(KC_ERR_INVALID, "n=%llu exceeds 32-bit vertex ids", (unsigned long long)n);
    if (n > 0 && (!csr->offsets || (!csr->targets && csr->offsets[n] > 0))) fail(KC_ERR_INVALID, "null CSR arrays");
    const uint64_t m2 = n ? csr->offsets[n] : 0;
    if (n && csr->offsets[0] != 0) fail(KC_ERR_INVALID, "offsets[0] must be 0");
    if (m2 >= (1ull << 32)) fail(KC_ERR_INVALID, "2m=%llu exceeds 2^32 directed edges", (unsigned long long)m2);
    uint64_t non_isolated = 0;
    for (uint64_t v = 0; v < n; ++v) {
      if (csr->offsets[v + 1] < csr->offsets[v]) fail(KC_ERR_INVALID, "offsets not monotone at %llu", (unsigned long long)v);
      non_isolated += csr->offsets[v + 1] > csr->offsets[v];
    }
    for (uint64_t e = 0; e < m2; ++e)
      if (csr->targets[e] >= n) fail(KC_ERR_INVALID, "target %u out of range at edge %llu", csr->targets[e], (unsigned long long)e);
    // the contraction reads every cluster edge from both of its ends
    uint64_t bu = 0, bv = 0;
    if (n && find_one_way_edge(csr->offsets, csr->targets, 0, n, bu, bv))
      fail(KC_ERR_INVALID, "CSR is not symmetric: edge %llu -> %llu has no reverse", (unsigned long long)bu,
           (unsigned long long)bv);
    std::unique_ptr<kc_graph> g = new_handle();
    g->n = n;
    g->m2 = m2;
    g->non_isolated = non_isolated;
    g->off0 = g->dalloc<uint64_t>(n + 1);
    g->col0 = g->dalloc<uint32_t>(m2);
    KC_HIP(hipMemcpy(g->off0, csr->offsets, (n + 1) * 8, hipMemcpyHostToDevice));
    if (m2) KC_HIP(hipMemcpy(g->col0, csr->targets, m2 * 4, hipMemcpyHostToDevice));
    init_work(g.get());
    *out = g.release();
  });
}

int kc_graph_generate(const char *name, const kc_gen_params *pp, kc_graph **out) {
  return guarded([&] {
    if (!name || !pp || !out) fail(KC_ERR_INVALID, "null argument");
    *out = nullptr;
    kc_gen_params p = *pp;
    const std::string gname(name);
    uint64_t n = 0, slots = 0;
    int dims = 0;
    uint32_t scale = 0;
    if (gname == "gnm_undirected") {
      n = p.n;
      slots = p.m;
    } else if (gname == "rmat") {
      while ((1ull << scale) < p.n) ++scale;
      if ((1ull << scale) != p.n || scale > 31) fail(KC_ERR_INVALID, "rmat: n=%llu must be a power of two below 2^32", (unsigned long long)p.n);
      n = p.n;
      slots = p.m;
    } else if (gname == "grid_2d" || gname == "grid_3d") {
      dims = gname == "grid_2d" ? 2 : 3;
      if (!p.nx) p.nx = p.n;
      if (!p.ny) p.ny = dims == 2 ? p.m : p.nx;
      if (dims == 3 && !p.nz) p.nz = p.nx;
      if (dims == 2) p.nz = 1;
      n = p.nx * p.ny * p.nz;
      slots = n * dims;
    } else if (gname == "ba") {
      n = p.n;
      slots = p.n * (p.d >= 1 ? static_cast<uint64_t>(p.d) : 1);
    } else {
      fail(KC_ERR_UNSUPPORTED, "device generator not available for %s (use kc_generate)", name);
    }
    if (n >= 0xFFFFFFFFull) fail(KC_ERR_INVALID, "%s: n=%llu exceeds 32-bit vertex ids", name, (unsigned long long)n);
    std::unique_ptr<kc_graph> g = new_handle();
    hipStream_t s = g->stream;
    {
      DevTemp src(slots * 4 + 64), dst(slots * 4 + 64);
      uint64_t count = slots;
      if (gname == "gnm_undirected") {
        gen_gnm(slots, n, p.seed, src.as<uint32_t>(), dst.as<uint32_t>(), s);
      } else if (gname == "rmat") {
        gen_rmat(slots, scale, p.rmat_a, p.rmat_b, p.rmat_c, p.seed, src.as<uint32_t>(), dst.as<uint32_t>(), s);
      } else if (gname == "ba") {
        gen_ba(slots, p.d >= 1 ? static_cast<uint64_t>(p.d) : 1, p.seed, src.as<uint32_t>(), dst.as<uint32_t>(), s);
      } else {
        if (p.nx >= (1ull << 32) || p.ny >= (1ull << 32) || p.nz >= (1ull << 32))
          fail(KC_ERR_INVALID, "%s: extents must be below 2^32", name);
        gen_grid(p.nx, p.ny, p.nz, dims, p.periodic, p.p, p.seed, src.as<uint32_t>(), dst.as<uint32_t>(), s);
      }
      g->n = n;
      build_device_csr(g.get(), n, src.as<uint32_t>(), dst.as<uint32_t>(), count);
    }
    init_work(g.get());
    *out = g.release();
  });
}

int kc_graph_read_binary(const char *path, kc_graph **out) {
  return guarded([&] {
    if (!path || !out) fail(KC_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<FILE, int (*)(FILE *)> fp(std::fopen(path, "rb"), std::fclose);
    if (!fp) fail(KC_ERR_IO, "Error opening %s", path);
    uint64_t head[2] = {0, 0};
    if (std::fread(head, sizeof(uint64_t), 2, fp.get()) != 2) fail(KC_ERR_IO, "%s: missing header (n and m as u64)", path);
    const uint64_t n = head[0], m = head[1];
    if (n >= 0xFFFFFFFFull) fail(KC_ERR_INVALID, "%s: n=%llu exceeds 32-bit vertex ids", path, (unsigned long long)n);
    std::fseek(fp.get(), 0, SEEK_END);
    const long size = std::ftell(fp.get());
    std::fseek(fp.get(), 2 * sizeof(uint64_t), SEEK_SET);
    const uint64_t present = size > 16 ? static_cast<uint64_t>(size - 16) / 16 : 0;
    if (present < m)
      fail(KC_ERR_IO, "%s: %llu of %llu edges present", path, (unsigned long long)present, (unsigned long long)m);
    std::unique_ptr<kc_graph> g = new_handle();
    hipStream_t s = g->stream;
    {
      // the pairs go through a pinned staging buffer to the device in chunks,
      // where they become 0-based u32 edges; the CSR is then built in HBM
      constexpr uint64_t kChunk = 1ull << 22;  // pairs per chunk (64 MB)
      DevTemp src(m * 4 + 64), dst(m * 4 + 64), raw(kChunk * 16), bad(64);
      void *pinned = nullptr;
      KC_HIP(hipHostMalloc(&pinned, kChunk * 16));
      std::unique_ptr<void, hipError_t (*)(void *)> pin(pinned, hipHostFree);
      fill_u32(bad.as<uint32_t>(), 1, 0, s);
      for (uint64_t done = 0; done < m;) {
        const uint64_t take = std::min<uint64_t>(kChunk, m - done);
        if (std::fread(pinned, 16, take, fp.get()) != take) fail(KC_ERR_IO, "%s: read failed", path);
        KC_HIP(hipMemcpyAsync(raw.p, pinned, take * 16, hipMemcpyHostToDevice, s));
        binary_edges(raw.as<uint64_t>(), take, n, src.as<uint32_t>() + done, dst.as<uint32_t>() + done,
                     bad.as<int>(), s);
        KC_HIP(hipStreamSynchronize(s));  // the staging buffer is refilled next
        done += take;
      }
      int h_bad = 0;
      KC_HIP(hipMemcpy(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost));
      if (h_bad) fail(KC_ERR_INVALID, "%s: an edge lies outside [1,%llu]", path, (unsigned long long)n);
      g->n = n;
      build_device_csr(g.get(), n, src.as<uint32_t>(), dst.as<uint32_t>(), m);
    }
    init_work(g.get());
    *out = g.release();
  });
}

}  // extern "C"

namespace {

// Edges of one sorted edge file, parsed on the device.
struct EdgeTextPart {
  std::unique_ptr<DevTemp> src, dst;
  uint64_t count = 0;
};

// Next whitespace-separated number of [p, e) (spaces and tabs only), as the
// host reader takes it; false when missing or malformed.
bool header_number(const char *&p, const char *e, unsigned long long &v) {
  while (p < e && (*p == ' ' || *p == '\t')) ++p;
  if (p >= e || *p < '0' || *p > '9') return false;
  unsigned long long x = 0;
  for (; p < e && *p >= '0' && *p <= '9'; ++p) {
    const unsigned d = static_cast<unsigned>(*p - '0');
    if (x > (~0ull - d) / 10) return false;
    x = x * 10 + d;
  }
  v = x;
  return true;
}

// ReadSortedEdgeFile (lib/io/graph_io.h:385-669): header "<type> n m" (checked
// against n and m when check is set), then the "<type> u v" lines, at most m
// of them when bounded.  The file goes to HBM in one copy and is parsed there.
void device_edge_file(kc_graph *g, const char *path, bool bounded, bool check, uint64_t &n, uint64_t &m,
                      EdgeTextPart &out) {
  std::unique_ptr<FILE, int (*)(FILE *)> fp(std::fopen(path, "rb"), std::fclose);
  if (!fp) fail(KC_ERR_IO, "Error opening %s", path);
  std::fseek(fp.get(), 0, SEEK_END);
  const long size = std::ftell(fp.get());
  std::fseek(fp.get(), 0, SEEK_SET);
  if (size < 0) fail(KC_ERR_IO, "%s: cannot size the file", path);
  const uint64_t fsize = static_cast<uint64_t>(size);
  void *pinned = nullptr;
  KC_HIP(hipHostMalloc(&pinned, fsize + 1));
  std::unique_ptr<void, hipError_t (*)(void *)> pin(pinned, hipHostFree);
  char *text = static_cast<char *>(pinned);
  if (fsize && std::fread(text, 1, fsize, fp.get()) != fsize) fail(KC_ERR_IO, "%s: read failed", path);
  // header: the first line that is empty or does not start with '%'
  uint64_t p = 0, hs = 0, he = 0;
  bool have = false;
  while (p < fsize) {
    const char *nl = static_cast<const char *>(std::memchr(text + p, '\n', fsize - p));
    uint64_t e = nl ? static_cast<uint64_t>(nl - text) : fsize;
    const uint64_t next = nl ? e + 1 : fsize;
    if (e > p && text[e - 1] == '\r') --e;
    if (e == p || text[p] != '%') {
      hs = p;
      he = e;
      have = true;
      p = next;
      break;
    }
    p = next;
  }
  if (!have) fail(KC_ERR_IO, "%s: missing header", path);
  const char *h = text + hs, *he_p = text + he;
  while (h < he_p && (*h == ' ' || *h == '\t')) ++h;
  while (h < he_p && *h != ' ' && *h != '\t') ++h;
  unsigned long long hn = 0, hm = 0;
  if (!header_number(h, he_p, hn) || !header_number(h, he_p, hm))
    fail(KC_ERR_IO, "%s: malformed header '%.*s' (expected '<type> n m')", path, static_cast<int>(he - hs), text + hs);
  if (hn >= 0xFFFFFFFFull) fail(KC_ERR_INVALID, "%s: n=%llu exceeds 32-bit vertex ids", path, hn);
  if (check && (hn != n || hm != m))
    fail(KC_ERR_INVALID, "%s: header n=%llu m=%llu differs from the first part's n=%llu m=%llu", path, hn, hm,
         static_cast<unsigned long long>(n), static_cast<unsigned long long>(m));
  n = hn;
  m = hm;
  const uint64_t bytes = fsize - p;
  hipStream_t s = g->stream;
  DevTemp body(bytes + 64), scratch((scan_scratch_words(bytes + 1) + 1024) * 8), total(64), bad(64);
  const uint8_t *b = body.as<uint8_t>();
  if (bytes) KC_HIP(hipMemcpyAsync(body.p, text + p, bytes, hipMemcpyHostToDevice, s));
  uint64_t *sc = scratch.as<uint64_t>(), *tot = total.as<uint64_t>();
  const uint64_t nl = byte_positions(b, bytes, kBytesNewline, nullptr, tot, sc, s);
  DevTemp nlpos(nl * 8 + 64);
  byte_positions(b, bytes, kBytesNewline, nlpos.as<uint64_t>(), tot, sc, s);
  const uint64_t lines = nl + (bytes && text[fsize - 1] != '\n' ? 1 : 0);
  uint64_t count = edge_text_lines(b, bytes, nlpos.as<uint64_t>(), nl, lines, nullptr, tot, sc, s);
  DevTemp eline(count * 8 + 64);
  edge_text_lines(b, bytes, nlpos.as<uint64_t>(), nl, lines, eline.as<uint64_t>(), tot, sc, s);
  if (bounded && count > m) count = m;
  out.src = std::make_unique<DevTemp>(count * 4 + 64);
  out.dst = std::make_unique<DevTemp>(count * 4 + 64);
  out.count = count;
  fill_u32(bad.as<uint32_t>(), 1, 0, s);
  edge_text(b, bytes, nlpos.as<uint64_t>(), nl, eline.as<uint64_t>(), count, n, out.src->as<uint32_t>(),
            out.dst->as<uint32_t>(), bad.as<int>(), s);
  int h_bad = 0;
  KC_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  KC_HIP(hipStreamSynchronize(s));
  if (h_bad & 1) fail(KC_ERR_IO, "%s: malformed edge line (expected '<type> u v')", path);
  if (h_bad & 2) fail(KC_ERR_INVALID, "%s: an edge lies outside [1,%llu]", path, static_cast<unsigned long long>(n));
}

// The edges of all parts in one list, then the CSR in HBM.
void device_edge_graph(kc_graph *g, uint64_t n, std::vector<EdgeTextPart> &parts) {
  uint64_t total = 0;
  for (const EdgeTextPart &e : parts) total += e.count;
  DevTemp src(total * 4 + 64), dst(total * 4 + 64);
  uint64_t at = 0;
  for (const EdgeTextPart &e : parts) {
    if (e.count) {
      KC_HIP(hipMemcpyAsync(src.as<uint32_t>() + at, e.src->p, e.count * 4, hipMemcpyDeviceToDevice, g->stream));
      KC_HIP(hipMemcpyAsync(dst.as<uint32_t>() + at, e.dst->p, e.count * 4, hipMemcpyDeviceToDevice, g->stream));
    }
    at += e.count;
  }
  parts.clear();
  g->n = n;
  build_device_csr(g, n, src.as<uint32_t>(), dst.as<uint32_t>(), total);
}

}  // namespace

extern "C" {

int kc_graph_read_edgefile(const char *path, kc_graph **out) {
  return guarded([&] {
    if (!path || !out) fail(KC_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<kc_graph> g = new_handle();
    uint64_t n = 0, m = 0;
    std::vector<EdgeTextPart> parts(1);
    device_edge_file(g.get(), path, true, false, n, m, parts[0]);
    device_edge_graph(g.get(), n, parts);
    init_work(g.get());
    *out = g.release();
  });
}

int kc_graph_read_partitioned_edgefile(const char *path, int parts_n, kc_graph **out) {
  return guarded([&] {
    if (!path || !out || parts_n < 1) fail(KC_ERR_INVALID, "null argument or parts < 1");
    *out = nullptr;
    std::unique_ptr<kc_graph> g = new_handle();
    uint64_t n = 0, m = 0;
    std::vector<EdgeTextPart> parts(static_cast<size_t>(parts_n));
    for (int r = 0; r < parts_n; ++r) {
      const std::string part = std::string(path) + "_" + std::to_string(r);
      device_edge_file(g.get(), part.c_str(), false, r > 0, n, m, parts[static_cast<size_t>(r)]);
    }
    device_edge_graph(g.get(), n, parts);
    init_work(g.get());
    *out = g.release();
  });
}

int kc_graph_read_metis(const char *path, kc_graph **out) {
  return guarded([&] {
    if (!path || !out) fail(KC_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<FILE, int (*)(FILE *)> fp(std::fopen(path, "rb"), std::fclose);
    if (!fp) fail(KC_ERR_IO, "Error opening %s", path);
    std::fseek(fp.get(), 0, SEEK_END);
    const long size = std::ftell(fp.get());
    std::fseek(fp.get(), 0, SEEK_SET);
    if (size < 0) fail(KC_ERR_IO, "%s: cannot size the file", path);
    const uint64_t fsize = static_cast<uint64_t>(size);
    // the whole file in pinned memory: one copy to HBM
    void *pinned = nullptr;
    KC_HIP(hipHostMalloc(&pinned, fsize + 1));
    std::unique_ptr<void, hipError_t (*)(void *)> pin(pinned, hipHostFree);
    char *text = static_cast<char *>(pinned);
    if (fsize && std::fread(text, 1, fsize, fp.get()) != fsize) fail(KC_ERR_IO, "%s: read failed", path);
    text[fsize] = '\0';
    // header: the first line not starting with '%' (its first number is n)
    uint64_t p = 0, hs = 0, he = 0;
    bool have = false;
    while (p < fsize) {
      const char *nl = static_cast<const char *>(std::memchr(text + p, '\n', fsize - p));
      const uint64_t e = nl ? static_cast<uint64_t>(nl - text) : fsize;
      if (text[p] != '%' || e == p) {
        hs = p;
        he = e;
        have = true;
        p = nl ? e + 1 : fsize;
        break;
      }
      p = nl ? e + 1 : fsize;
    }
    if (!have) fail(KC_ERR_IO, "%s: missing header", path);
    const std::string header(text + hs, text + he);
    char *q = nullptr;
    errno = 0;
    const unsigned long long n = std::strtoull(header.c_str(), &q, 10);
    if (q == header.c_str() || errno) fail(KC_ERR_IO, "%s: malformed header '%s'", path, header.c_str());
    if (n >= 0xFFFFFFFFull) fail(KC_ERR_INVALID, "%s: n=%llu exceeds 32-bit vertex ids", path, n);
    const uint64_t bytes = fsize - p;
    std::unique_ptr<kc_graph> g = new_handle();
    hipStream_t s = g->stream;
    {
      DevTemp body(bytes + 64), scratch((scan_scratch_words(bytes + 1) + 1024) * 8), total(64), bad(64);
      const uint8_t *b = body.as<uint8_t>();
      if (bytes) KC_HIP(hipMemcpyAsync(body.p, text + p, bytes, hipMemcpyHostToDevice, s));
      uint64_t *sc = scratch.as<uint64_t>(), *tot = total.as<uint64_t>();
      // lines: positions of the newlines; a last line without one counts too
      const uint64_t nl = byte_positions(b, bytes, kBytesNewline, nullptr, tot, sc, s);
      DevTemp nlpos(nl * 8 + 64);
      byte_positions(b, bytes, kBytesNewline, nlpos.as<uint64_t>(), tot, sc, s);
      const uint64_t lines = nl + (bytes && text[fsize - 1] != '\n' ? 1 : 0);
      DevTemp vline(lines * 4 + 64);
      metis_vertex_lines(b, bytes, nlpos.as<uint64_t>(), lines, n, vline.as<uint32_t>(), tot, sc, s);
      fill_u32(vline.as<uint32_t>() + lines, 1, kNone, s);  // past the last line (not reached by a number)
      // numbers, and bytes that are neither digits nor blanks
      const uint64_t count = byte_positions(b, bytes, kBytesToken, nullptr, tot, sc, s);
      const uint64_t other = byte_positions(b, bytes, kBytesOther, nullptr, tot, sc, s);
      fill_u32(bad.as<uint32_t>(), 1, 0, s);
      if (other) {
        DevTemp opos(other * 8 + 64);
        byte_positions(b, bytes, kBytesOther, opos.as<uint64_t>(), tot, sc, s);
        metis_other(opos.as<uint64_t>(), other, nlpos.as<uint64_t>(), nl, vline.as<uint32_t>(), bad.as<int>(), s);
        KC_HIP(hipStreamSynchronize(s));
      }
      DevTemp tokpos(count * 8 + 64), tv(count * 4 + 64), tval(count * 4 + 64);
      byte_positions(b, bytes, kBytesToken, tokpos.as<uint64_t>(), tot, sc, s);
      metis_tokens(b, bytes, tokpos.as<uint64_t>(), count, nlpos.as<uint64_t>(), nl, vline.as<uint32_t>(), n,
                   tv.as<uint32_t>(), tval.as<uint32_t>(), bad.as<int>(), s);
      int h_bad = 0;
      KC_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
      KC_HIP(hipStreamSynchronize(s));
      if (h_bad & 2) fail(KC_ERR_IO, "%s: a vertex line holds a character other than digits and blanks", path);
      if (h_bad & 1) fail(KC_ERR_INVALID, "%s: a neighbour id is 0 or outside [1,%llu]", path, n);
      DevTemp src(count * 4 + 64), dst(count * 4 + 64);
      const uint64_t edges = metis_edges(count, tv.as<uint32_t>(), tval.as<uint32_t>(), src.as<uint32_t>(),
                                         dst.as<uint32_t>(), tot, sc, s);
      g->n = n;
      build_device_csr(g.get(), n, src.as<uint32_t>(), dst.as<uint32_t>(), edges);
    }
    init_work(g.get());
    *out = g.release();
  });
}

}  // extern "C"

namespace kc {

kc_graph *graph_device_begin(uint64_t n, uint64_t m2, uint64_t **off, uint32_t **col, hipStream_t *stream) {
  if (n >= 0xFFFFFFFFull) fail(KC_ERR_INVALID, "n=%llu exceeds 32-bit vertex ids", (unsigned long long)n);
  if (m2 >= (1ull << 32)) fail(KC_ERR_INVALID, "2m=%llu exceeds 2^32 directed edges", (unsigned long long)m2);
  std::unique_ptr<kc_graph> g = new_handle();
  g->n = n;
  g->m2 = m2;
  g->off0 = g->dalloc<uint64_t>(n + 1);
  g->col0 = g->dalloc<uint32_t>(m2);
  *off = g->off0;
  *col = g->col0;
  *stream = g->stream;
  return g.release();
}

void graph_device_finish(kc_graph *g) {
  hipStream_t s = g->stream;
  unsigned long long *cnt = reinterpret_cast<unsigned long long *>(g->sizes + 6);
  fill_u64(g->sizes + 6, 1, 0, s);
  nonzero_rows(g->n, g->off0, cnt, s);
  KC_HIP(hipMemcpyAsync(g->h_sizes + 6, g->sizes + 6, 8, hipMemcpyDeviceToHost, s));
  KC_HIP(hipStreamSynchronize(s));
  g->non_isolated = g->h_sizes[6];
  init_work(g);
}

const uint32_t *graph_device_labels(const kc_graph *g) { return g->labels; }

}  // namespace kc

extern "C" {

int kc_graph_download(const kc_graph *g, kc_csr *out) {
  return guarded([&] {
    if (!g || !out) fail(KC_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->n = g->n;
    out->m = g->m2 / 2;
    out->offsets = static_cast<uint64_t *>(std::malloc((g->n + 1) * 8));
    out->targets = static_cast<uint32_t *>(std::malloc(g->m2 ? g->m2 * 4 : 4));
    if (!out->offsets || !out->targets) fail(KC_ERR_OOM, "host allocation failed");
    KC_HIP(hipMemcpy(out->offsets, g->off0, (g->n + 1) * 8, hipMemcpyDeviceToHost));
    if (g->m2) KC_HIP(hipMemcpy(out->targets, g->col0, g->m2 * 4, hipMemcpyDeviceToHost));
  });
}

int kc_graph_destroy(kc_graph *g) {
  return guarded([&] { delete g; });
}

int kc_graph_info(const kc_graph *g, uint64_t *n, uint64_t *m, uint64_t *non_isolated) {
  return guarded([&] {
    if (!g) fail(KC_ERR_INVALID, "null graph");
    if (n) *n = g->n;
    if (m) *m = g->m2 / 2;
    if (non_isolated) *non_isolated = g->non_isolated;
  });
}

int kc_find_components(kc_graph *g, const kc_options *opt, uint32_t *labels_out, kc_stats *stats) {
  return guarded([&] {
    if (!g) fail(KC_ERR_INVALID, "null graph");
    kc_options o;
    if (opt)
      o = *opt;
    else
      kc_options_init(&o);
    if (o.variant < KC_VARIANT_EXP || o.variant > KC_VARIANT_LABELS) fail(KC_ERR_INVALID, "unknown variant %d", o.variant);
    if (o.dedup_huge < KC_DEDUP_AUTO || o.dedup_huge > KC_DEDUP_MAPS_TABLES)
      fail(KC_ERR_INVALID, "unknown dedup_huge mode %d", o.dedup_huge);
    if (o.time_class < 0 || o.time_class > KC_KERNEL_CLASSES)
      fail(KC_ERR_INVALID, "time_class must be in [0, %d], got %d", KC_KERNEL_CLASSES, o.time_class);
    if (o.pre_contract < KC_PRE_NONE || o.pre_contract > KC_PRE_DOUBLE)
      fail(KC_ERR_INVALID, "unknown pre_contract mode %d", o.pre_contract);
    if (o.pre_contract != KC_PRE_NONE) {
      if (o.variant != KC_VARIANT_EXP && o.variant != KC_VARIANT_LOCAL)
        fail(KC_ERR_UNSUPPORTED, "pre-contraction applies to the contraction variants only");
      if (o.pre_blocks < 1) fail(KC_ERR_INVALID, "pre_blocks must be >= 1 with pre_contract %d", o.pre_contract);
    }
    if (o.replicate) {
      if (o.variant != KC_VARIANT_EXP) fail(KC_ERR_UNSUPPORTED, "replicate applies to exponential contraction only");
      if (o.pre_contract != KC_PRE_NONE) fail(KC_ERR_UNSUPPORTED, "replicate does not combine with pre_contract");
      if (!(o.hd_threshold >= 0.0) || !(o.hd_threshold < 1e12)) fail(KC_ERR_INVALID, "hd_threshold must be in [0, 1e12)");
    }
    if (o.learned) {
      if (o.variant != KC_VARIANT_EXP) fail(KC_ERR_UNSUPPORTED, "learned centre proposals apply to exponential contraction only");
      if (!g->mlp) fail(KC_ERR_INVALID, "learned centre proposals need a model (kc_graph_set_mlp)");
      if (!(o.learned_alpha >= 0.f) || o.learned_alpha > 64.f) fail(KC_ERR_INVALID, "learned_alpha must be in [0, 64]");
      ensure_learned_buffers(g);
    }
    const auto t0 = std::chrono::steady_clock::now();
    g->have_labels = false;
    if (o.variant == KC_VARIANT_EXP) ensure_table(g, o.beta);
    g->arena.reset();
    g->timer.begin_run(o.time_kernels != 0, g->stream, o.time_class - 1);
    KC_HIP(hipEventRecord(g->ev_begin, g->stream));
    RunOut out;
    if (o.variant == KC_VARIANT_EXP || o.variant == KC_VARIANT_LOCAL)
      run_contraction(g, &o, out);
    else
      run_propagation(g, &o, out);
    KC_HIP(hipEventRecord(g->ev_end, g->stream));
    d2h(g, g->h_small + 5, g->small + 5, sizeof(uint32_t));
    sync(g);
    g->components = g->n ? g->h_small[5] : 0;
    if (o.variant == KC_VARIANT_EXP || o.variant == KC_VARIANT_LOCAL)
      g->components += static_cast<uint32_t>(g->n - g->n0);  // degree-0 vertices
    g->have_labels = true;
    float ms = 0.f;
    KC_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
    if (labels_out && g->n) KC_HIP(hipMemcpy(labels_out, g->labels, g->n * 4, hipMemcpyDeviceToHost));
    const auto t1 = std::chrono::steady_clock::now();
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->time_ms = ms;
      stats->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
      stats->n = g->n;
      stats->m = g->m2 / 2;
      stats->levels = out.levels;
      stats->components = g->components;
      stats->sweeps = out.sweeps;
      stats->radius_sum = out.radius_sum;
      stats->replicas = out.replicas;
      stats->hubs = out.hubs;
      const BinCfg cfg = bin_cfg(g, &o);
      stats->bin_small = cfg.tmax;
      stats->bin_mid = cfg.smax;
      stats->bin_large = cfg.wmax;
      g->timer.finish(stats);
      uint64_t tot = 0;
      for (int i = 0; i < KC_KERNEL_CLASSES; ++i) tot += stats->kernel_bytes[i];
      stats->alg_bytes = tot;
    }
  });
}

int kc_graph_set_bins(kc_graph *g, const kc_bin_tuning *t) {
  return guarded([&] {
    if (!g) fail(KC_ERR_INVALID, "null graph");
    if (!t) {
      g->tuned = BinCfg{0, 0, 0};
      return;
    }
    if (t->small <= 0 || t->mid <= 0 || t->large <= 0) fail(KC_ERR_INVALID, "bin thresholds must be positive");
    g->tuned = bin_config(t->small, t->mid, t->large);
  });
}

int kc_graph_autotune(kc_graph *g, const kc_options *opt, kc_bin_tuning *chosen) {
  return guarded([&] {
    if (!g) fail(KC_ERR_INVALID, "null graph");
    kc_options o;
    if (opt)
      o = *opt;
    else
      kc_options_init(&o);
    o.time_kernels = 0;
    o.record_trace = 0;
    // one lane, 16 lanes, one wave (and the hub chunk size): the defaults and
    // their neighbours by factors of 2 to 4
    const int32_t smalls[] = {8, 16, 32}, mids[] = {128, 256, 512}, larges[] = {1024, 4096, 16384};
    auto time_of = [&](int32_t a, int32_t b, int32_t c) {
      kc_options t = o;
      t.sweep_group_small = a;
      t.sweep_group_mid = b;
      t.sweep_group_large = c;
      kc_stats st;
      double sum = 0;
      for (int rep = 0; rep < 3; ++rep) {
        t.seed = o.seed + rep;
        const int rc = kc_find_components(g, &t, nullptr, &st);
        if (rc != KC_OK) throw Failure{rc};
        if (rep) sum += st.time_ms;  // the first call warms the bins and buffers
      }
      return sum / 2;
    };
    const double def_ms = time_of(16, 256, 4096);
    double best = def_ms;
    int32_t bs = 16, bm = 256, bl = 4096, n = 1;
    for (int32_t a : smalls)
      for (int32_t b : mids)
        for (int32_t c : larges) {
          if (a == 16 && b == 256 && c == 4096) continue;
          const double ms = time_of(a, b, c);
          ++n;
          if (ms < best) {
            best = ms;
            bs = a;
            bm = b;
            bl = c;
          }
        }
    // a choice within 2% of the defaults is run-to-run noise: keep the defaults
    if (best > 0.98 * def_ms) {
      best = def_ms;
      bs = 16;
      bm = 256;
      bl = 4096;
    }
    g->tuned = bin_config(bs, bm, bl);
    if (chosen) {
      chosen->small = static_cast<int32_t>(g->tuned.tmax);
      chosen->mid = static_cast<int32_t>(g->tuned.smax);
      chosen->large = static_cast<int32_t>(g->tuned.wmax);
      chosen->ms = best;
      chosen->default_ms = def_ms;
      chosen->candidates = n;
    }
  });
}

int kc_component_histogram(kc_graph *g, uint64_t *pairs, uint64_t capacity, uint64_t *num_pairs) {
  return guarded([&] {
    if (!g || !num_pairs) fail(KC_ERR_INVALID, "null argument");
    if (!g->have_labels) fail(KC_ERR_INVALID, "no labels: kc_find_components has not completed");
    const uint64_t n = g->n;
    // distinct sizes k satisfy k (k + 1) / 2 <= n
    uint64_t kmax = 1;
    while (kmax * (kmax + 1) / 2 <= n) ++kmax;
    uint32_t *hist = g->arena.alloc_n<uint32_t>(n + 1);
    uint64_t *dpairs = g->arena.alloc_n<uint64_t>(2 * kmax + 2);
    component_histogram(n, g->labels, g->minid, hist, dpairs, g->sizes + 10, g->scan_scratch, g->stream);
    d2h(g, g->h_sizes + 10, g->sizes + 10, sizeof(uint64_t));
    sync(g);
    const uint64_t np = g->h_sizes[10];
    if (np > kmax) fail(KC_ERR_INVALID, "internal: %llu component sizes", (unsigned long long)np);
    *num_pairs = np;
    if (pairs && capacity) {
      const uint64_t take = np < capacity ? np : capacity;
      if (take) KC_HIP(hipMemcpy(pairs, dpairs, 16 * take, hipMemcpyDeviceToHost));
    }
  });
}

int kc_graph_set_mlp(kc_graph *g, kc_mlp *m) {
  return guarded([&] {
    if (!g) fail(KC_ERR_INVALID, "null graph");
    if (m) KC_HIP(hipStreamSynchronize(m->stream));
    g->mlp = m;
    g->l0_score_mlp = nullptr;  // level 0's scores are computed again for the new model
  });
}

int kc_graph_features(kc_graph *g, float *features) {
  return guarded([&] {
    if (!g || !features) fail(KC_ERR_INVALID, "null argument");
    if (!g->n0) return;
    ensure_learned_buffers(g);
    level_degrees(g->n0, g->off0c, g->ldeg, g->stream);
    level_features(g->n0, g->off0c, g->col0c, g->ldeg, 0, g->feat, g->stream);
    KC_HIP(hipMemcpyAsync(features, g->feat, g->n0 * kMlpIn * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    sync(g);
  });
}

int kc_mlp_collect(kc_mlp *m, kc_graph *g, const kc_options *opt, uint64_t stride) {
  return guarded([&] {
    if (!m || !g) fail(KC_ERR_INVALID, "null argument");
    kc_options o;
    if (opt)
      o = *opt;
    else
      kc_options_init(&o);
    if (o.variant != KC_VARIANT_EXP || o.learned) fail(KC_ERR_INVALID, "traces come from the plain exponential contraction");
    ensure_learned_buffers(g);
    KC_HIP(hipStreamSynchronize(m->stream));
    g->collect = m;
    g->collect_stride = stride ? stride : 1;
    struct Reset {
      kc_graph *g;
      ~Reset() { g->collect = nullptr; }
    } reset{g};
    const int rc = kc_find_components(g, &o, nullptr, nullptr);
    if (rc != KC_OK) throw Failure{rc};
    sync(g);
  });
}

int kc_copy_labels(kc_graph *g, uint32_t *labels_out) {
  return guarded([&] {
    if (!g || !labels_out) fail(KC_ERR_INVALID, "null argument");
    if (!g->have_labels) fail(KC_ERR_INVALID, "no labels: kc_find_components has not completed");
    if (g->n) KC_HIP(hipMemcpy(labels_out, g->labels, g->n * 4, hipMemcpyDeviceToHost));
  });
}

int kc_trace_levels(const kc_graph *g, uint32_t *levels) {
  return guarded([&] {
    if (!g || !levels) fail(KC_ERR_INVALID, "null argument");
    *levels = static_cast<uint32_t>(g->levels.size());
  });
}

int kc_trace_level_info(const kc_graph *g, uint32_t level, uint64_t info[4]) {
  return guarded([&] {
    if (!g || !info) fail(KC_ERR_INVALID, "null argument");
    if (level >= g->levels.size()) fail(KC_ERR_INVALID, "level %u out of range", level);
    const LevelRec &r = g->levels[level];
    info[0] = r.n;
    info[1] = r.m2 / 2;
    info[2] = r.sweeps;
    info[3] = r.radius;
  });
}

int kc_trace_level_keys(const kc_graph *g, uint32_t level, uint64_t *keys_out, uint32_t *gid_out) {
  return guarded([&] {
    if (!g) fail(KC_ERR_INVALID, "null graph");
    if (level >= g->levels.size()) fail(KC_ERR_INVALID, "level %u out of range", level);
    const LevelRec &r = g->levels[level];
    if (keys_out) {
      if (!r.keys_copy) fail(KC_ERR_INVALID, "run without record_trace");
      KC_HIP(hipMemcpy(keys_out, r.keys_copy, r.n * 8, hipMemcpyDeviceToHost));
      if (r.idx) {  // stored in another order: back to level-index order
        std::vector<uint64_t> byp(keys_out, keys_out + r.n);
        std::vector<uint32_t> idx(r.n);
        KC_HIP(hipMemcpy(idx.data(), r.idx, r.n * 4, hipMemcpyDeviceToHost));
        for (uint64_t p = 0; p < r.n; ++p) keys_out[idx[p]] = byp[p];
      }
    }
    if (gid_out) {
      if (r.gid) {
        KC_HIP(hipMemcpy(gid_out, r.gid, r.n * 4, hipMemcpyDeviceToHost));
      } else {
        for (uint64_t v = 0; v < r.n; ++v) gid_out[v] = static_cast<uint32_t>(v);
      }
    }
  });
}

int kc_deviates(uint32_t gid0, uint32_t count, int32_t seed, uint32_t iteration, uint64_t rng_offset, double beta,
                int32_t fma, int32_t stream, uint32_t *out) {
  return guarded([&] {
    if (!out && count) fail(KC_ERR_INVALID, "null output");
    if (!(beta > 1e-6) || !(beta < 1e6)) fail(KC_ERR_INVALID, "beta must be in (1e-6, 1e6), got %g", beta);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) fail(KC_ERR_HIP, "no HIP device available");
    std::vector<double> t = build_deviate_table(beta);
    double *dt = nullptr;
    uint32_t *dout = nullptr;
    KC_HIP(hipMalloc(&dt, (t.size() + 1) * 8));
    KC_HIP(hipMalloc(&dout, (static_cast<size_t>(count) + 1) * 4));
    KC_HIP(hipMemcpy(dt, t.data(), t.size() * 8, hipMemcpyHostToDevice));
    DeviateArgs d{seed, iteration, rng_offset, fma, stream, dt, static_cast<uint32_t>(t.size())};
    deviates_only(gid0, count, d, dout, nullptr);
    KC_HIP(hipDeviceSynchronize());
    KC_HIP(hipMemcpy(out, dout, static_cast<size_t>(count) * 4, hipMemcpyDeviceToHost));
    (void)hipFree(dt);
    (void)hipFree(dout);
  });
}

}  // extern "C"
