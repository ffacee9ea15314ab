// Mixed batches of small graphs for the batched form: per graph a size drawn
// log-uniformly, an average degree drawn uniformly and a type drawn uniformly
// among G(n,m), 2D random geometric and bond-percolated 2D/3D grids.  Every
// draw is a pure function of (seed, graph index), so the batch does not depend
// on the thread count.  Graphs are generated in parallel (one graph per
// thread) with the single-graph generators and packed as CSR segments with
// local ids.
#include <omp.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gen_rng.h"
#include "kacomp.h"

namespace kc {
void set_error(const char *fmt, ...);
}

extern "C" void kc_batch_gen_params_init(kc_batch_gen_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->graphs = 4096;
  p->log_n_min = 8;
  p->log_n_max = 18;
  p->deg_min = 2;
  p->deg_max = 16;
  p->types = 7;
  p->seed = 4;
}

extern "C" void kc_batch_free(kc_batch_csr *b) {
  if (!b) return;
  std::free(b->vtx_off);
  std::free(b->edge_off);
  std::free(b->offsets);
  std::free(b->targets);
  std::memset(b, 0, sizeof(*b));
}

extern "C" int kc_batch_generate(const kc_batch_gen_params *pp, kc_batch_csr *out) {
  if (!pp || !out) {
    kc::set_error("null argument");
    return KC_ERR_INVALID;
  }
  std::memset(out, 0, sizeof(*out));
  const kc_batch_gen_params p = *pp;
  int kinds[3], nk = 0;
  for (int t = 0; t < 3; ++t)
    if (p.types & (1 << t)) kinds[nk++] = t;
  if (nk == 0 || p.graphs == 0 || !(p.log_n_min >= 0) || p.log_n_max < p.log_n_min || p.log_n_max > 30 ||
      !(p.deg_min >= 0) || p.deg_max < p.deg_min) {
    kc::set_error("batch generator: bad parameters (graphs %u, log n [%g, %g], degree [%g, %g], types %d)",
                  p.graphs, p.log_n_min, p.log_n_max, p.deg_min, p.deg_max, p.types);
    return KC_ERR_INVALID;
  }
  const kc::GenStream rs = kc::gen_stream(p.seed, 7);
  std::vector<kc_csr> parts(p.graphs);
  std::vector<int> rc(p.graphs, KC_OK);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t gi = 0; gi < static_cast<int64_t>(p.graphs); ++gi) {
    omp_set_num_threads(1);  // the single-graph generator runs on this thread only
    const uint64_t i = static_cast<uint64_t>(gi);
    const double u1 = kc::gen_unit(kc::gen_value(rs, 4 * i));
    const double u2 = kc::gen_unit(kc::gen_value(rs, 4 * i + 1));
    const uint64_t z3 = kc::gen_value(rs, 4 * i + 2);
    const uint64_t z4 = kc::gen_value(rs, 4 * i + 3);
    const uint64_t n = static_cast<uint64_t>(std::llround(std::exp2(p.log_n_min + (p.log_n_max - p.log_n_min) * u1)));
    const double d = p.deg_min + (p.deg_max - p.deg_min) * u2;
    const int kind = kinds[kc::gen_bounded(z3, static_cast<uint64_t>(nk))];
    kc_gen_params g;
    kc_gen_params_init(&g);
    g.seed = static_cast<int32_t>(static_cast<uint32_t>(p.seed) * 1000003u + static_cast<uint32_t>(i));
    std::string name;
    if (kind == 0) {
      name = "gnm_undirected";
      g.n = n;
      g.m = static_cast<uint64_t>(std::llround(static_cast<double>(n) * d / 2.0));
    } else if (kind == 1) {
      name = "rgg_2d";
      g.n = n;
      g.r = std::sqrt(d / (3.14159265358979323846 * static_cast<double>(n)));
    } else {
      const int dims = (z4 & 1) ? 3 : 2;
      name = dims == 2 ? "grid_2d" : "grid_3d";
      uint64_t side = static_cast<uint64_t>(std::llround(std::pow(static_cast<double>(n), 1.0 / dims)));
      if (side < 2) side = 2;
      g.nx = side;
      g.ny = side;
      g.nz = dims == 3 ? side : 1;
      g.n = side;
      g.m = side;
      g.p = std::fmin(1.0, d / (2.0 * dims));
    }
    rc[gi] = kc_generate(name.c_str(), &g, &parts[gi]);
  }
  for (uint32_t g = 0; g < p.graphs; ++g)
    if (rc[g] != KC_OK) {
      for (auto &c : parts) kc_csr_free(&c);
      return rc[g];
    }
  out->graphs = p.graphs;
  out->vtx_off = static_cast<uint64_t *>(std::malloc(8ull * (p.graphs + 1)));
  out->edge_off = static_cast<uint64_t *>(std::malloc(8ull * (p.graphs + 1)));
  if (!out->vtx_off || !out->edge_off) {
    for (auto &c : parts) kc_csr_free(&c);
    kc_batch_free(out);
    kc::set_error("host allocation failed");
    return KC_ERR_OOM;
  }
  out->vtx_off[0] = out->edge_off[0] = 0;
  for (uint32_t g = 0; g < p.graphs; ++g) {
    out->vtx_off[g + 1] = out->vtx_off[g] + parts[g].n;
    out->edge_off[g + 1] = out->edge_off[g] + parts[g].offsets[parts[g].n];
  }
  const uint64_t n = out->vtx_off[p.graphs], m2 = out->edge_off[p.graphs];
  out->offsets = static_cast<uint64_t *>(std::malloc(8ull * (n + 1)));
  out->targets = static_cast<uint32_t *>(std::malloc(4ull * (m2 ? m2 : 1)));
  if (!out->offsets || !out->targets) {
    for (auto &c : parts) kc_csr_free(&c);
    kc_batch_free(out);
    kc::set_error("host allocation failed");
    return KC_ERR_OOM;
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t gi = 0; gi < static_cast<int64_t>(p.graphs); ++gi) {
    kc_csr &c = parts[gi];
    const uint64_t v0 = out->vtx_off[gi], e0 = out->edge_off[gi];
    for (uint64_t v = 0; v < c.n; ++v) out->offsets[v0 + v] = e0 + c.offsets[v];
    if (c.offsets[c.n]) std::memcpy(out->targets + e0, c.targets, 4ull * c.offsets[c.n]);
    kc_csr_free(&c);
  }
  out->offsets[n] = m2;
  return KC_OK;
}
