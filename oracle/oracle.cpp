// CPU restatement of the connected-components contraction path of
// sebalamm/components, in the limit of one vertex per PE (every vertex is an
// interface vertex, so the local-component pre-contraction of
// CAGBuilder::BuildLocalComponentGraph keeps the whole graph and exponential
// contraction runs on it).  TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Compiled with -ffp-contract=off: every fused multiply-add is explicit.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <queue>
#include <utility>
#include <vector>

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint64_t kHop = 1ull << 32;

struct Csr {
  uint64_t n = 0;
  std::vector<uint64_t> off;
  std::vector<uint32_t> tgt;
};

// Union-find with union by smaller root id and path halving; the root of each
// set is its minimum member, which is the canonical label.
struct MinUnionFind {
  std::vector<uint32_t> p;
  explicit MinUnionFind(uint64_t n) : p(n) {
    for (uint64_t i = 0; i < n; ++i) p[i] = static_cast<uint32_t>(i);
  }
  uint32_t find(uint32_t x) {
    while (p[x] != x) {
      p[x] = p[p[x]];
      x = p[x];
    }
    return x;
  }
  void unite(uint32_t a, uint32_t b) {
    a = find(a);
    b = find(b);
    if (a == b) return;
    if (a < b)
      p[b] = a;
    else
      p[a] = b;
  }
};

void canonical(uint64_t n, const uint64_t *off, const uint32_t *tgt, uint32_t *labels) {
  MinUnionFind uf(n);
  for (uint64_t v = 0; v < n; ++v)
    for (uint64_t e = off[v]; e < off[v + 1]; ++e) uf.unite(static_cast<uint32_t>(v), tgt[e]);
  for (uint64_t v = 0; v < n; ++v) labels[v] = uf.find(static_cast<uint32_t>(v));
}

// exponential_contraction.h:467-471, restated: minstd_rand0 seeded with
// (unsigned)(gid + seed + iteration*(rng_offset + seed)), two outputs combined
// as in generate_canonical<double,53> (C++11 [rand.util.canonical]: k = 2 calls
// of 30 bits for R = 2^31-2), then -log(1-u)/beta truncated.
uint32_t exp_deviate(uint32_t gid, int32_t seed, uint32_t iteration, uint64_t rng_offset, double beta, bool fma) {
  const uint64_t g = gid;
  const uint64_t s64 = g + static_cast<uint64_t>(static_cast<int64_t>(seed)) +
                       static_cast<uint64_t>(iteration) * (rng_offset + static_cast<uint64_t>(static_cast<int64_t>(seed)));
  const uint32_t s = static_cast<uint32_t>(s64);
  const uint64_t m = 2147483647ull;
  uint64_t x = s % m;
  if (x == 0) x = 1;
  x = (x * 16807ull) % m;
  const double g1 = static_cast<double>(x - 1);
  x = (x * 16807ull) % m;
  const double g2 = static_cast<double>(x - 1);
  const double r = static_cast<double>(m - 1);  // max - min + 1
  double sum = 0.0;
  double tmp = 1.0;
  sum = sum + g1 * tmp;
  tmp = tmp * r;
  if (fma)
    sum = std::fma(g2, tmp, sum);
  else
    sum = sum + g2 * tmp;
  tmp = tmp * r;
  double u = sum / tmp;
  if (u >= 1.0) u = std::nextafter(1.0, 0.0);
  const double dev = -std::log(1.0 - u) / beta;
  return static_cast<uint32_t>(static_cast<uint64_t>(1.0 * dev));
}

// Hashed stream of the library (SplitMix64 over seed, iteration, gid), then
// the exponential law evaluated directly with std::log.  The law is the one
// of exponential_contraction.h:467-471 (exponential_distribution(beta),
// truncated); the stream is not the reference's minstd_rand0 draw.
// SplitMix64's output function (Steele, Lea, Flood 2014).
uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double hashed_uniform(uint32_t gid, int32_t seed, uint32_t iteration) {
  const uint64_t key = ((static_cast<uint64_t>(static_cast<uint32_t>(seed)) << 32) | iteration) * 0x9e3779b97f4a7c15ull;
  const uint64_t z = mix64(mix64(key) + static_cast<uint64_t>(gid) * 0xd1b54a32d192ed03ull);
  return static_cast<double>(z >> 11) * 0x1p-53;
}
uint32_t hashed_deviate(uint32_t gid, int32_t seed, uint32_t iteration, double beta) {
  const double u = hashed_uniform(gid, seed, iteration);
  return static_cast<uint32_t>(static_cast<uint64_t>(-std::log(1.0 - u) / beta));
}

// Uniform deviate in [0, 99] of local contraction: the law of
// uniform_int_distribution<unsigned>(0, 99) (local_contraction.h:173-176),
// drawn from a SplitMix64 stream over (seed, iteration, gid) and reduced by a
// multiply-shift of its upper 32 bits.
uint32_t uniform100(uint32_t gid, int32_t seed, uint32_t iteration) {
  const uint64_t z = mix64((0x9e3779b97f4a7c15ull * (static_cast<uint64_t>(static_cast<uint32_t>(seed)) + 1)) ^
                           (static_cast<uint64_t>(iteration) << 32) ^ gid);
  return static_cast<uint32_t>(((z >> 32) * 100ull) >> 32);
}

// The LP sweep loop, in place, vertex order (one PE holds every vertex).
uint64_t lp_fixpoint(uint64_t n, const uint64_t *off, const uint32_t *tgt, std::vector<uint64_t> &key, uint64_t hop) {
  uint64_t rounds = 0;
  bool converged = false;
  while (!converged) {
    converged = true;
    for (uint64_t v = 0; v < n; ++v) {
      uint64_t best = key[v];
      for (uint64_t e = off[v]; e < off[v + 1]; ++e) {
        const uint64_t cand = key[tgt[e]] + hop;
        if (cand < best) {
          best = cand;
          converged = false;
        }
      }
      key[v] = best;
    }
    ++rounds;
  }
  return rounds;
}

void lp_closed_form(uint64_t n, const uint64_t *off, const uint32_t *tgt, const uint64_t *key0, uint64_t hop,
                    uint64_t *out) {
  std::vector<uint64_t> dist(key0, key0 + n);
  using Item = std::pair<uint64_t, uint32_t>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  for (uint64_t v = 0; v < n; ++v) pq.push({dist[v], static_cast<uint32_t>(v)});
  while (!pq.empty()) {
    auto [d, v] = pq.top();
    pq.pop();
    if (d != dist[v]) continue;
    for (uint64_t e = off[v]; e < off[v + 1]; ++e) {
      const uint32_t w = tgt[e];
      if (d + hop < dist[w]) {
        dist[w] = d + hop;
        pq.push({dist[w], w});
      }
    }
  }
  std::copy(dist.begin(), dist.end(), out);
}

}  // namespace

struct LevelTrace {
  uint64_t n = 0, m = 0, sweeps = 0;
  uint32_t radius = 0;
  std::vector<uint64_t> keys;
  std::vector<uint32_t> gid, lab, newidx;
};

struct or_result {
  std::vector<uint32_t> labels;
  std::vector<LevelTrace> levels;
  uint64_t sweeps = 0;
  double seconds = 0.0;
  uint64_t hubs = 0, replicas = 0;
};

namespace {

// Contracted graph of one level (DynamicContraction::ExponentialContraction,
// dynamic_contraction.h:62-136 with FindExponentialConflictingEdges :251-290
// and InsertAndClearEdges :513-559): vertices are the centres that keep an
// inter-cluster edge, edges the distinct centre pairs.
void contract(const Csr &g, const std::vector<uint32_t> &lab, std::vector<uint32_t> &newidx, Csr &next) {
  std::vector<uint64_t> pairs;
  for (uint64_t v = 0; v < g.n; ++v)
    for (uint64_t e = g.off[v]; e < g.off[v + 1]; ++e) {
      const uint32_t a = lab[v], b = lab[g.tgt[e]];
      if (a != b) pairs.push_back((static_cast<uint64_t>(a) << 32) | b);
    }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  newidx.assign(g.n, kNone);
  uint32_t k = 0;
  for (size_t i = 0; i < pairs.size(); ++i) {
    const uint32_t a = static_cast<uint32_t>(pairs[i] >> 32);
    if (newidx[a] == kNone) newidx[a] = k++;
  }
  next.n = k;
  next.off.assign(k + 1, 0);
  next.tgt.resize(pairs.size());
  for (size_t i = 0; i < pairs.size(); ++i) {
    const uint32_t a = newidx[pairs[i] >> 32];
    next.off[a + 1]++;
    next.tgt[i] = newidx[static_cast<uint32_t>(pairs[i])];
  }
  for (uint32_t a = 0; a < k; ++a) next.off[a + 1] += next.off[a];
}

// Local components of a vertex-block distribution (FindLocalComponents,
// exponential_contraction.h:392-411, with PE p holding the ids
// [p*bs, (p+1)*bs)): the components of the edges inside a block, each
// represented by its smallest member, as the reference's local BFS started
// from its vertices in id order.  Interface vertices (a neighbour in another
// block) keep themselves when only_inner is set: the default branch's
// BuildLocalComponentGraph (cag_builder.h:160-230) contracts the others only;
// -single's BuildComponentAdjacencyGraph contracts every vertex.
std::vector<uint32_t> block_components(const Csr &g, uint64_t blocks, bool only_inner) {
  const uint64_t bs = (g.n + blocks - 1) / blocks;
  std::vector<uint32_t> parent(g.n);
  for (uint64_t v = 0; v < g.n; ++v) parent[v] = static_cast<uint32_t>(v);
  auto find = [&](uint32_t v) {
    while (parent[v] != v) v = parent[v] = parent[parent[v]];
    return v;
  };
  std::vector<char> inner(g.n, 1);
  for (uint64_t v = 0; v < g.n; ++v)
    for (uint64_t e = g.off[v]; e < g.off[v + 1]; ++e) {
      const uint32_t u = g.tgt[e];
      if (u / bs != v / bs) {
        inner[v] = 0;
        continue;
      }
      const uint32_t a = find(static_cast<uint32_t>(v)), b = find(u);
      if (a != b) parent[std::max(a, b)] = std::min(a, b);  // the root stays the smallest member
    }
  std::vector<uint32_t> lab(g.n);
  for (uint64_t v = 0; v < g.n; ++v) lab[v] = only_inner && !inner[v] ? static_cast<uint32_t>(v) : find(static_cast<uint32_t>(v));
  return lab;
}

// Proxy slot in [0, k) of the edge (v, w) of a hub v: SplitMix64 over the two
// ids and the seed, multiply-shift reduced (the library's replica_pick).
uint32_t proxy_slot(uint32_t gv, uint32_t gw, int32_t seed, uint32_t k) {
  const uint64_t z = mix64((static_cast<uint64_t>(gv) << 32 | gw) ^
                           (0x9e3779b97f4a7c15ull * (static_cast<uint64_t>(static_cast<uint32_t>(seed)) + 1)));
  return static_cast<uint32_t>(((z >> 32) * k) >> 32);
}

// DistributeHighDegreeVertices (exponential_contraction.h:935-1253) with every
// vertex its own PE, so that every edge of a hub is remote: a hub (degree >=
// threshold * sqrt(n), utils.h:108-117) of degree d has floor(sqrt(d)) proxy
// slots (:1040, min(P - 1, sqrt(d)) with P = n); each of its edges moves to the
// slot proxy_slot names, an edge between two hubs joins the two proxies, and
// the slots that receive an edge become proxies n, n + 1, ... (hub order, then
// slot order), each joined to its hub, which keeps only its proxies.  A proxy
// takes its hub's id (its deviate is the hub's, :483-503).
void replicate_level0(Csr &cur, std::vector<uint32_t> &gid, int32_t seed, double threshold, or_result &res) {
  const uint64_t n = cur.n;
  if (!n) return;
  const uint64_t limit = std::max<uint64_t>(1, static_cast<uint64_t>(threshold * std::sqrt(static_cast<double>(n))));
  std::vector<uint32_t> hub_of(n, kNone), hubs, k;
  std::vector<uint64_t> base(1, 0);
  for (uint64_t v = 0; v < n; ++v)
    if (cur.off[v + 1] - cur.off[v] >= limit) {
      hub_of[v] = static_cast<uint32_t>(hubs.size());
      hubs.push_back(static_cast<uint32_t>(v));
      const uint64_t d = cur.off[v + 1] - cur.off[v];
      uint32_t kk = static_cast<uint32_t>(std::min<uint64_t>(n - 1, static_cast<uint64_t>(std::sqrt(static_cast<double>(d)))));
      k.push_back(kk ? kk : 1);
      base.push_back(base.back() + k.back());
    }
  if (hubs.empty()) return;
  std::vector<uint32_t> cnt(base.back(), 0);
  for (size_t h = 0; h < hubs.size(); ++h) {
    const uint32_t v = hubs[h];
    for (uint64_t e = cur.off[v]; e < cur.off[v + 1]; ++e) cnt[base[h] + proxy_slot(gid[v], gid[cur.tgt[e]], seed, k[h])]++;
  }
  std::vector<uint32_t> proxy(base.back(), kNone), first(hubs.size()), count(hubs.size(), 0), root;
  for (size_t h = 0; h < hubs.size(); ++h) {
    first[h] = static_cast<uint32_t>(n + root.size());
    for (uint64_t sl = base[h]; sl < base[h + 1]; ++sl)
      if (cnt[sl]) {
        proxy[sl] = static_cast<uint32_t>(n + root.size());
        root.push_back(hubs[h]);
        ++count[h];
      }
  }
  auto far_end = [&](uint32_t u, uint32_t w) {  // w as seen from u's edge
    const uint32_t hw = hub_of[w];
    return hw == kNone ? w : proxy[base[hw] + proxy_slot(gid[w], gid[u], seed, k[hw])];
  };
  const uint64_t R = root.size();
  std::vector<std::vector<uint32_t>> prow(R);
  for (uint64_t p = 0; p < R; ++p) prow[p].push_back(root[p]);
  for (size_t h = 0; h < hubs.size(); ++h) {
    const uint32_t v = hubs[h];
    for (uint64_t e = cur.off[v]; e < cur.off[v + 1]; ++e) {
      const uint32_t w = cur.tgt[e];
      prow[proxy[base[h] + proxy_slot(gid[v], gid[w], seed, k[h])] - n].push_back(far_end(v, w));
    }
  }
  Csr out;
  out.n = n + R;
  out.off.assign(out.n + 1, 0);
  for (uint64_t v = 0; v < n; ++v) {
    const uint32_t h = hub_of[v];
    if (h != kNone) {
      for (uint32_t j = 0; j < count[h]; ++j) out.tgt.push_back(first[h] + j);
    } else {
      for (uint64_t e = cur.off[v]; e < cur.off[v + 1]; ++e) out.tgt.push_back(far_end(static_cast<uint32_t>(v), cur.tgt[e]));
    }
    out.off[v + 1] = out.tgt.size();
  }
  for (uint64_t p = 0; p < R; ++p) {
    out.tgt.insert(out.tgt.end(), prow[p].begin(), prow[p].end());
    out.off[n + p + 1] = out.tgt.size();
    gid.push_back(gid[root[p]]);
  }
  cur = std::move(out);
  res.hubs = hubs.size();
  res.replicas = R;
}

void run_contraction(const Csr &g0, int variant, int32_t seed, double beta, bool fma, int stream,
                     uint64_t seq_limit, bool keep_keys, int pre_mode, uint64_t pre_blocks, bool replicate,
                     double hd_threshold, or_result &res) {
  // Level 0 holds the vertices of degree > 0 in id order: the reference sets
  // degree-0 vertices inactive at load (lib/io/graph_io.h:380) and its vertex
  // loops skip inactive vertices (lib/data_structures/dynamic_graph.h:113-117).
  std::vector<uint32_t> newid(g0.n, kNone), gid;
  for (uint64_t v = 0; v < g0.n; ++v)
    if (g0.off[v + 1] > g0.off[v]) {
      newid[v] = static_cast<uint32_t>(gid.size());
      gid.push_back(static_cast<uint32_t>(v));
    }
  const std::vector<uint32_t> gid0 = gid;
  const uint64_t rng_offset = gid.size();
  Csr cur;
  cur.n = gid.size();
  cur.off.assign(cur.n + 1, 0);
  cur.tgt.reserve(g0.off[g0.n]);
  for (uint64_t a = 0; a < cur.n; ++a) {
    for (uint64_t e = g0.off[gid[a]]; e < g0.off[gid[a] + 1]; ++e) cur.tgt.push_back(newid[g0.tgt[e]]);
    cur.off[a + 1] = cur.tgt.size();
  }
  if (replicate && variant == 0) replicate_level0(cur, gid, seed, hd_threshold, res);
  uint32_t iteration = 1;
  std::vector<uint32_t> top_rep;
  bool have_top = false;
  // pre-contraction levels: 1 local (default branch) or single, 2 for -contraction
  int pre_left = pre_mode == 3 ? 2 : pre_mode != 0 ? 1 : 0;
  while (cur.n > 0) {
    const uint64_t m2 = cur.off[cur.n];
    const bool pre = pre_left > 0 && m2 > 0;
    if (pre) --pre_left;
    if (!pre && (m2 == 0 || (seq_limit > 0 && m2 / 2 <= seq_limit))) {
      // RunSequentialCC (exponential_contraction.h:908-922): components of the
      // remaining graph, represented by their smallest member.
      std::vector<uint32_t> l(cur.n);
      canonical(cur.n, cur.off.data(), cur.tgt.data(), l.data());
      top_rep.resize(cur.n);
      for (uint64_t v = 0; v < cur.n; ++v) top_rep[v] = gid[l[v]];
      have_top = true;
      break;
    }
    LevelTrace t;
    t.n = cur.n;
    t.m = m2 / 2;
    std::vector<uint64_t> key(cur.n);
    if (pre) {
      // the clusters are the block components; distance 0 everywhere
      const std::vector<uint32_t> comp = block_components(cur, pre_blocks, pre_mode == 1);
      for (uint64_t v = 0; v < cur.n; ++v) key[v] = comp[v];
      t.sweeps = 0;
    } else if (variant == 1) {
      // LocalContraction::RunContraction (local_contraction.h:157-300): uniform
      // deviates, min over N[v], then min over N[N[v]], both synchronous.
      for (uint64_t v = 0; v < cur.n; ++v)
        key[v] = (static_cast<uint64_t>(uniform100(gid[v], seed, iteration)) << 32) | v;
      for (int pass = 0; pass < 2; ++pass) {
        std::vector<uint64_t> nk(key);
        for (uint64_t v = 0; v < cur.n; ++v)
          for (uint64_t e = cur.off[v]; e < cur.off[v + 1]; ++e) nk[v] = std::min(nk[v], key[cur.tgt[e]]);
        key.swap(nk);
      }
      t.sweeps = 2;
    } else {
      for (uint64_t v = 0; v < cur.n; ++v)
        key[v] = (static_cast<uint64_t>(stream == 0 ? exp_deviate(gid[v], seed, iteration, rng_offset, beta, fma)
                                                    : hashed_deviate(gid[v], seed, iteration, beta))
                  << 32) |
                 v;
      t.sweeps = lp_fixpoint(cur.n, cur.off.data(), cur.tgt.data(), key, kHop);
    }
    t.lab.resize(cur.n);
    for (uint64_t v = 0; v < cur.n; ++v) {
      t.lab[v] = static_cast<uint32_t>(key[v]);
      if (variant != 1 && !pre) {
        const uint32_t r = static_cast<uint32_t>(key[v] >> 32) - static_cast<uint32_t>(key[t.lab[v]] >> 32);
        t.radius = std::max(t.radius, r);
      }
    }
    Csr next;
    contract(cur, t.lab, t.newidx, next);
    std::vector<uint32_t> gid_next(next.n);
    for (uint64_t a = 0; a < cur.n; ++a)
      if (t.newidx[a] != kNone) gid_next[t.newidx[a]] = gid[a];
    t.gid = gid;
    if (keep_keys) t.keys = key;
    res.sweeps += t.sweeps;
    res.levels.push_back(std::move(t));
    cur = std::move(next);
    gid.swap(gid_next);
    if (!pre) ++iteration;
  }
  // UndoContraction (dynamic_contraction.h:568-658): every vertex takes the
  // representative of its centre one level up, or the centre itself when the
  // cluster was a whole component.
  std::vector<uint32_t> rep = top_rep;
  for (int l = static_cast<int>(res.levels.size()) - 1; l >= 0; --l) {
    const LevelTrace &t = res.levels[l];
    std::vector<uint32_t> down(t.n);
    for (uint64_t v = 0; v < t.n; ++v) {
      const uint32_t a = t.lab[v];
      down[v] = t.newidx[a] != kNone ? rep[t.newidx[a]] : t.gid[a];
    }
    rep.swap(down);
  }
  (void)have_top;
  // ApplyToLocalComponents with canonical (minimum id) labels; a degree-0
  // vertex is its own component.
  res.labels.resize(g0.n);
  for (uint64_t v = 0; v < g0.n; ++v) res.labels[v] = static_cast<uint32_t>(v);
  std::vector<uint32_t> minid(g0.n, kNone);
  for (uint64_t a = 0; a < gid0.size(); ++a) minid[rep[a]] = std::min(minid[rep[a]], gid0[a]);
  for (uint64_t a = 0; a < gid0.size(); ++a) res.labels[gid0[a]] = minid[rep[a]];
}

// Propagation (propagation.h:72-149) and ShortcutPropagation
// (shortcut_propagation.h:124-158): synchronous min-label rounds, the latter
// followed by one pointer jump l(v) <- min(l(v), l(l(v))) per round.
void run_propagation(const Csr &g, bool shortcut, or_result &res) {
  std::vector<uint32_t> label(g.n), next(g.n);
  for (uint64_t v = 0; v < g.n; ++v) label[v] = static_cast<uint32_t>(v);
  uint64_t rounds = 0;
  while (true) {
    for (uint64_t v = 0; v < g.n; ++v) {
      uint32_t best = label[v];
      for (uint64_t e = g.off[v]; e < g.off[v + 1]; ++e) best = std::min(best, label[g.tgt[e]]);
      next[v] = best;
    }
    if (shortcut)
      for (uint64_t v = 0; v < g.n; ++v) next[v] = std::min(next[v], label[next[v]]);
    ++rounds;
    if (next == label) break;
    label.swap(next);
  }
  res.sweeps = rounds;
  res.labels = label;
}

}  // namespace

extern "C" {

void or_canonical_labels(uint64_t n, const uint64_t *off, const uint32_t *tgt, uint32_t *labels) {
  canonical(n, off, tgt, labels);
}

uint32_t or_exp_deviate(uint32_t gid, int32_t seed, uint32_t iteration, uint64_t rng_offset, double beta, int32_t fma) {
  return exp_deviate(gid, seed, iteration, rng_offset, beta, fma != 0);
}

uint32_t or_hashed_deviate(uint32_t gid, int32_t seed, uint32_t iteration, double beta) {
  return hashed_deviate(gid, seed, iteration, beta);
}

uint32_t or_uniform_deviate_100(uint32_t gid, int32_t seed, uint32_t iteration) {
  return uniform100(gid, seed, iteration);
}

uint64_t or_splitmix64_mix(uint64_t z) { return mix64(z); }

double or_hashed_uniform(uint32_t gid, int32_t seed, uint32_t iteration) { return hashed_uniform(gid, seed, iteration); }

uint64_t or_lp_fixpoint(uint64_t n, const uint64_t *off, const uint32_t *tgt, const uint64_t *key0, uint64_t hop,
                        uint64_t *key_out) {
  std::vector<uint64_t> key(key0, key0 + n);
  uint64_t r = lp_fixpoint(n, off, tgt, key, hop);
  std::copy(key.begin(), key.end(), key_out);
  return r;
}

void or_lp_closed_form(uint64_t n, const uint64_t *off, const uint32_t *tgt, const uint64_t *key0, uint64_t hop,
                       uint64_t *key_out) {
  lp_closed_form(n, off, tgt, key0, hop, key_out);
}

or_result *or_run(uint64_t n, const uint64_t *off, const uint32_t *tgt, int32_t variant, int32_t seed, double beta,
                  int32_t fma, int32_t stream, uint64_t seq_limit, int32_t keep_keys, int32_t pre_mode,
                  uint64_t pre_blocks, int32_t replicate, double hd_threshold) {
  Csr g;
  g.n = n;
  g.off.assign(off, off + n + 1);
  g.tgt.assign(tgt, tgt + off[n]);
  or_result *r = new or_result();
  const auto t0 = std::chrono::steady_clock::now();
  if (variant == 0 || variant == 1)
    run_contraction(g, variant, seed, beta, fma != 0, stream, seq_limit, keep_keys != 0, pre_mode,
                    pre_blocks ? pre_blocks : 1, replicate != 0, hd_threshold, *r);
  else
    run_propagation(g, variant == 2, *r);
  r->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return r;
}

void or_free(or_result *r) { delete r; }
void or_labels(const or_result *r, uint32_t *labels) { std::copy(r->labels.begin(), r->labels.end(), labels); }
uint32_t or_levels(const or_result *r) { return static_cast<uint32_t>(r->levels.size()); }
void or_level_info(const or_result *r, uint32_t level, uint64_t info[4]) {
  const LevelTrace &t = r->levels[level];
  info[0] = t.n;
  info[1] = t.m;
  info[2] = t.sweeps;
  info[3] = t.radius;
}
void or_level_keys(const or_result *r, uint32_t level, uint64_t *keys, uint32_t *gid) {
  const LevelTrace &t = r->levels[level];
  if (keys) std::copy(t.keys.begin(), t.keys.end(), keys);
  if (gid) std::copy(t.gid.begin(), t.gid.end(), gid);
}
uint64_t or_total_sweeps(const or_result *r) { return r->sweeps; }
uint64_t or_replicate(uint64_t n, const uint64_t *off, const uint32_t *tgt, int32_t seed, double hd_threshold,
                      uint64_t *off_out, uint32_t *tgt_out, uint32_t *gid_out) {
  // level 0 (the vertices of degree > 0 in id order), then its hubs split
  std::vector<uint32_t> newid(n, kNone), gid;
  for (uint64_t v = 0; v < n; ++v)
    if (off[v + 1] > off[v]) {
      newid[v] = static_cast<uint32_t>(gid.size());
      gid.push_back(static_cast<uint32_t>(v));
    }
  Csr cur;
  cur.n = gid.size();
  cur.off.assign(cur.n + 1, 0);
  for (uint64_t a = 0; a < cur.n; ++a) {
    for (uint64_t e = off[gid[a]]; e < off[gid[a] + 1]; ++e) cur.tgt.push_back(newid[tgt[e]]);
    cur.off[a + 1] = cur.tgt.size();
  }
  or_result res;
  replicate_level0(cur, gid, seed, hd_threshold, res);
  if (off_out) {
    std::copy(cur.off.begin(), cur.off.end(), off_out);
    std::copy(cur.tgt.begin(), cur.tgt.end(), tgt_out);
    std::copy(gid.begin(), gid.end(), gid_out);
  }
  return cur.n;
}

void or_replication(const or_result *r, uint64_t out[2]) {
  out[0] = r->hubs;
  out[1] = r->replicas;
}
double or_seconds(const or_result *r) { return r->seconds; }

}  // extern "C"
