// Host CSR construction: degree count, scatter of both directions, per-row
// sort + unique, compaction.  Three passes over the edges, parallel over rows.
#include "csr_build.h"

#include <omp.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace kc {

bool alloc_csr(kc_csr *out, uint64_t n, uint64_t m2) {
  out->n = n;
  out->m = m2 / 2;
  out->offsets = static_cast<uint64_t *>(std::malloc((n + 1) * sizeof(uint64_t)));
  out->targets = static_cast<uint32_t *>(std::malloc((m2 ? m2 : 1) * sizeof(uint32_t)));
  if (!out->offsets || !out->targets) {
    std::free(out->offsets);
    std::free(out->targets);
    out->offsets = nullptr;
    out->targets = nullptr;
    return false;
  }
  return true;
}

void build_symmetric_csr(uint64_t n, const uint32_t *src, const uint32_t *dst, uint64_t count, kc_csr *out) {
  std::vector<uint64_t> deg(n + 1, 0);
  // degrees of both directions (self-loops dropped)
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(count); ++i) {
    const uint32_t u = src[i], v = dst[i];
    if (u == v) continue;
    __atomic_fetch_add(&deg[u], 1, __ATOMIC_RELAXED);
    __atomic_fetch_add(&deg[v], 1, __ATOMIC_RELAXED);
  }
  std::vector<uint64_t> off(n + 1, 0);
  for (uint64_t v = 0; v < n; ++v) off[v + 1] = off[v] + deg[v];
  const uint64_t total = off[n];
  std::vector<uint32_t> tgt(total ? total : 1);
  std::vector<uint64_t> pos(off.begin(), off.end() - 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(count); ++i) {
    const uint32_t u = src[i], v = dst[i];
    if (u == v) continue;
    tgt[__atomic_fetch_add(&pos[u], 1, __ATOMIC_RELAXED)] = v;
    tgt[__atomic_fetch_add(&pos[v], 1, __ATOMIC_RELAXED)] = u;
  }
  // sort + unique per row
  std::vector<uint64_t> udeg(n + 1, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < static_cast<int64_t>(n); ++v) {
    uint32_t *b = tgt.data() + off[v];
    uint32_t *e = tgt.data() + off[v + 1];
    std::sort(b, e);
    udeg[v] = static_cast<uint64_t>(std::unique(b, e) - b);
  }
  uint64_t m2 = 0;
  for (uint64_t v = 0; v < n; ++v) m2 += udeg[v];
  if (!alloc_csr(out, n, m2)) return;
  out->offsets[0] = 0;
  for (uint64_t v = 0; v < n; ++v) out->offsets[v + 1] = out->offsets[v] + udeg[v];
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < static_cast<int64_t>(n); ++v)
    std::memcpy(out->targets + out->offsets[v], tgt.data() + off[v], udeg[v] * sizeof(uint32_t));
}

bool find_one_way_edge(const uint64_t *offsets, const uint32_t *targets, uint64_t a, uint64_t z, uint64_t &u,
                       uint64_t &v) {
  const uint64_t rows = z - a, base = offsets[a], slots = offsets[z] - base;
  if (!slots) return false;
  const bool par = slots > (1u << 16);
  std::vector<uint32_t> sorted(targets + base, targets + base + slots);
#pragma omp parallel for schedule(dynamic, 1024) if (par)
  for (int64_t r = 0; r < static_cast<int64_t>(rows); ++r)
    std::sort(sorted.data() + (offsets[a + r] - base), sorted.data() + (offsets[a + r + 1] - base));
  uint64_t bad = ~0ull;  // smallest row holding a one-way edge
#pragma omp parallel for schedule(dynamic, 1024) reduction(min : bad) if (par)
  for (int64_t r = 0; r < static_cast<int64_t>(rows); ++r) {
    for (uint64_t e = offsets[a + r]; e < offsets[a + r + 1]; ++e) {
      const uint32_t w = targets[e];
      const uint32_t *b = sorted.data() + (offsets[a + w] - base), *end = sorted.data() + (offsets[a + w + 1] - base);
      if (!std::binary_search(b, end, static_cast<uint32_t>(r))) {
        bad = std::min<uint64_t>(bad, static_cast<uint64_t>(r));
        break;
      }
    }
  }
  if (bad == ~0ull) return false;
  u = bad;
  for (uint64_t e = offsets[a + bad]; e < offsets[a + bad + 1]; ++e) {
    const uint32_t w = targets[e];
    const uint32_t *b = sorted.data() + (offsets[a + w] - base), *end = sorted.data() + (offsets[a + w + 1] - base);
    if (!std::binary_search(b, end, static_cast<uint32_t>(bad))) {
      v = w;
      break;
    }
  }
  return true;
}

bool find_one_way_edge_batch(uint32_t graphs, const uint64_t *vtx_off, const uint64_t *offsets,
                             const uint32_t *targets, uint32_t &graph, uint64_t &u, uint64_t &v) {
  uint32_t bad = ~0u;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : bad)
  for (int64_t g = 0; g < static_cast<int64_t>(graphs); ++g) {
    uint64_t a = 0, b = 0;
    if (find_one_way_edge(offsets, targets, vtx_off[g], vtx_off[g + 1], a, b))
      bad = std::min(bad, static_cast<uint32_t>(g));
  }
  if (bad == ~0u) return false;
  graph = bad;
  return find_one_way_edge(offsets, targets, vtx_off[bad], vtx_off[bad + 1], u, v);
}

}  // namespace kc
